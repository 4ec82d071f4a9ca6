// Source sampling, 1D-spherical IMC/DDMC particle transport and the flux tally on gfx950.
//
//   k_sources    SOURCE/interior_source.f90:57-168 + source_transformdirection.f90:23-46
//   k_transport  particle_advance.f90:140-294 with TRANSPORT1/transport11.f90:86-489
//                (one IMC event) and TRANSPORT1/diffusion11.f90:91-631 (one DDMC event)
//   k_fluxtally  fluxtally.f:30-54
//
// Transport is a persistent kernel over SoA particle arrays: a lane runs one particle to
// escape or death, then takes the next from a global counter, so lanes stay busy until
// the pool is drained; every particle is capped at max_steps events.  The event count of
// SuperLite's random walk (no census: a particle leaves only by escaping or dying) is
// dominated by effective scatters, so the group samplers they call are the exact bucket
// maps of slgpu_sample.h: one table load for most draws instead of nested searches.
// Per-particle randomness is a Philox-4x32-10 stream keyed on (seed, iteration) and
// counted on (particle, step, phase, draw within the step), so trajectories do not depend
// on the schedule, and every lane of a pass computes its step's first block together.
//
// Tallies: the per-cell energy tally lives in LDS (nc doubles) per block and is flushed
// with FP64 atomics at the end; a particle's consecutive events in one cell are merged in
// registers before the LDS atomic.  Each event adds to grd_jrad with one FP64 atomic into
// one of its replicas (block-indexed), which k_jrad_reduce folds together afterwards.
#include "slgpu_device.h"
#include "slgpu_internal.h"
#include "slgpu_sample.h"

// No fused multiply-add contraction in this file: the reference evaluates each product and
// sum separately, and a contraction can turn an exact cancellation into a negative
// rounding residue (the grazing shell-crossing root sqrt(x_l^2 - (1-mu^2)x^2) at x = x_l
// becomes the square root of -ulp, a NaN path length).
#pragma clang fp contract(off)

namespace slgpu {

#define CAPT(ig, ic) ((double)d.cap32[(size_t)((ic) - 1) * ng + ((ig) - 1)])
#define XA(i) (d.xarr[(i) - 1])
#define VXA(i) (d.vxarr[(i) - 1])
#define DXX(i) (d.xarr[(i)] - d.xarr[(i) - 1])
#define DX3X(i) (d.xarr[(i)] * d.xarr[(i)] * d.xarr[(i)] - d.xarr[(i) - 1] * d.xarr[(i) - 1] * d.xarr[(i) - 1])
#define SIGT(ic) (d.sig[(ic) - 1])
#define OPL(k, ic) (d.opaclump[(size_t)((ic) - 1) * 11 + ((k) - 1)])
#define WLG(ig) (d.wl[(ig) - 1])
#define WLI(ig) (d.wlinv[(ig) - 1])
#define DIVV(ix) (d.divv[(ix) - 1])
#define TINV(ic) (d.tempinv[(ic) - 1])

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_sources(Dev d, uint32_t iteration)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.npart) return;
    const int ng = d.ng, nx = d.nx;
    // owning cell: particles are numbered cell by cell (interior_source.f90:60-88)
    int lo = 0, hi = nx;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (d.pstart[mid] <= p) lo = mid; else hi = mid;
    }
    const int i = lo + 1;
    const int64_t nemit = d.pstart[i] - d.pstart[i - 1];
    Philox rng;
    rng.init(d.seed, iteration, (uint32_t)p, 0u);
    double r1 = rng.draw();
    const double *cum = d.srccum + (size_t)(i - 1) * ng;
    int iig = first_above(cum, 0, ng, r1) + 1;
    if (iig > ng) iig = ng;
    r1 = rng.draw();
    const double wl0 = 1.0 / ((1.0 - r1) * WLI(iig) + r1 * WLI(iig + 1));
    const double ep0 = d.emit[i - 1] / (double)nemit;
    r1 = rng.draw();
    const double mu0 = sqrt(r1);
    r1 = rng.draw();       // azimuth (om), unused in 1D spherical transport
    const int l1 = i - 1 < 1 ? 1 : i - 1, l2 = i + 1 > nx ? nx : i + 1;
    double uul = 0.5 * (d.emit[l1 - 1] + d.emit[i - 1]);
    double uur = 0.5 * (d.emit[l2 - 1] + d.emit[i - 1]);
    const double uumax = uul > uur ? uul : uur;
    uul = uul / uumax;
    uur = uur / uumax;
    const double third = (double)(1.0f / 3.0f);   // single-precision 1.0/3.0 of the reference
    const double xl = XA(i), xr = XA(i + 1);
    const double xl3 = xl * xl * xl, xr3 = xr * xr * xr;
    double r2 = 1.0, r3 = 0.0, x0 = xl;
    int guard = 0;
    while (r2 > r3 && guard < (1 << 20)) {
        r1 = rng.draw();
        x0 = pow(r1 * xr3 + (1.0 - r1) * xl3, third);
        r3 = (x0 - xl) / DXX(i);
        r3 = r3 * uur + (1.0 - r3) * uul;
        r2 = rng.draw();
        ++guard;
    }
    int ix = binsrch(x0, d.xarr, nx + 1, false);
    ix = ix < 1 ? i : ix;
    const double h1 = XA(ix), h2 = XA(ix + 1);
    const double vx = (VXA(ix) * (h2 - x0) + VXA(ix + 1) * (x0 - h1)) / (h2 - h1);
    const double cmffact = 1.0 + mu0 * vx * CINV;
    double mu = (mu0 + vx * CINV) / cmffact;
    mu = mu > 1.0 ? 1.0 : (mu < -1.0 ? -1.0 : mu);
    d.px[p] = x0;
    d.pmu[p] = mu;
    d.pe[p] = ep0;
    d.pe0[p] = ep0;
    d.pwl[p] = wl0;
    if (guard >= (1 << 20)) atomicOr(d.flags, 1);
}

// ---------------------------------------------------------------------------
struct Ptcl {
    double x, mu, e, e0, wl, vx;
    int ix, ig, itype, stat, idist;
    // for a DDMC particle: NF_VALID | BT_LUMP of its current (group, cell) when the step that
    // put it there knew them, so its next step needs no opacity load to pick its thresholds
    uint32_t nflag;
};
constexpr uint32_t NF_VALID = 1u;
// A DDMC -> IMC conversion (diffusion11.f90:329-355, 402-423, 515-538, 587-618) is recorded
// in nflag and completed at the particle's next pass, an IMC pass, where most lanes are
// such conversions: bits 1-3 the kind, bit 4 set when the step's draw 2 was still unused.
// The adiabatic factor of the converting step and that draw wait in the block's LDS.
enum { PEND_SCAT = 1, PEND_LEFT = 2, PEND_RIGHT = 3, PEND_DOP = 4 };
constexpr uint32_t NF_PEND = 7u << 1;
constexpr uint32_t NF_R2FREE = 1u << 4;

// TRANSPORT1/transport11.f90:86-489 (trn_noampfact = .true.)
// Arithmetic of the precision form.  SLGPU_MIXED evaluates the sampled path lengths and
// event times (-ln r), the adiabatic factor and the r^3-uniform position in FP32; the
// shell-crossing quadratic, the Newton-Raphson Doppler distance, the DDMC thresholds,
// positions, energies and every tally stay FP64.
template <bool M> __device__ __forceinline__ double p_log(double x)
{
    if constexpr (M) return (double)__logf((float)x);
    else return log(x);
}
// -ln(r)/k: a sampled path length for opacity k (transport11.f90:112-125)
template <bool M> __device__ __forceinline__ double p_path(double r, double k)
{
    if constexpr (M) return -(double)__logf((float)r) / k;
    else return -log(r) / k;
}
// exp(-divv*t/3): the DDMC adiabatic factor (diffusion11.f90:623-629).  A particle takes
// ~1e4-1e5 of these factors, so a systematic last-bit error in an FP32 factor near 1
// compounds into a percent-level energy bias.  The precision form therefore evaluates the
// loss fraction d = 1 - exp(-a) in FP32 to FP32 *relative* accuracy (series for small a)
// and forms the factor 1 - d in FP64.
template <bool M> __device__ __forceinline__ double p_adiabatic(double divv, double t)
{
    if constexpr (M) {
        const float a = (float)(divv * t) * (1.0f / 3.0f);
        const float d = a < 0.03f
                            ? a * (1.0f - a * (0.5f - a * (1.0f / 6.0f - a * (1.0f / 24.0f))))
                            : 1.0f - expf(-a);
        return 1.0 - (double)d;
    } else {
        return exp(-divv * t / 3);
    }
}
// r^3-uniform radius in [xl, xr]: (r*xr^3 + (1-r)*xl^3)**(1.0/3.0) with the reference's
// single-precision exponent e = (1.0/3.0 in single) = (1 + 2^-25)/3 (diffusion11.f90:336,597).
// Radii cubed (~1e47 cm^3) overflow FP32, so the precision form takes the FP32 root of the
// normalised b = r + (1-r)*(xl/xr)^3 and multiplies by xr^(3e) = xr * exp(2^-25 * ln xr).
// That last factor (~1 + 1.1e-6 at supernova radii) is what carries the outermost draws past
// xr into the caller's clamp, as in the FP64 form; dropping it moves ~0.1% of the
// conversions in thin shells off the face and biased the step count by +0.75%.
template <bool M> __device__ __forceinline__ double p_shell_root(double r1, double xl, double xr)
{
    if constexpr (M) {
        const double q = xl / xr;
        const double dl = 0x1p-25 * (double)__logf((float)xr);
        const double xr3e = xr * (1.0 + dl * (1.0 + 0.5 * dl));
        return xr3e * (double)powf((float)(r1 + (1.0 - r1) * (q * q * q)), 1.0f / 3.0f);
    } else {
        return pow(r1 * (xr * xr * xr) + (1.0 - r1) * (xl * xl * xl), (double)(1.0f / 3.0f));
    }
}

// transport11.f90:259-285: the experimental anisotropic Thomson cosine for draw r1.  The
// second cube root has a negative base for r1 below ~0.957, where the real power is NaN;
// every range test then fails and the isotropic cosine is taken, as in the reference.
__device__ __forceinline__ double thomson_aniso_mu(double r1)
{
    const double q = sqrt(1.0 + 16.0 * ((r1 - 0.5) * (r1 - 0.5)));
    const double mutemp = pow(4.0 * r1 - 2.0 + q, 1.0 / 3.0) - pow(q - (4.0 - 2.0 * r1), 1.0 / 3.0);
    const double muhelp = mutemp;
    const double B = 2 * mutemp * muhelp;
    const double C = mutemp * mutemp + muhelp * muhelp - 1;
    const double munew1 = (-B + sqrt(B * B - 4 * C)) / 2;
    const double munew2 = (-B - sqrt(B * B - 4 * C)) / 2;
    const bool ok1 = munew1 <= 1 && munew1 >= -1, ok2 = munew2 <= 1 && munew2 >= -1;
    if (ok1 && ok2) return r1 < 0.5 ? munew1 : munew2;
    if (ok1) return munew1;
    if (ok2) return munew2;
    return 1.0 - 2.0 * r1;
}

template <bool M>
__device__ int transport11(const Dev &d, Ptcl &p, StepRng &s, double &erd, double &jrd,
                           double &eamp, double &evelo)
{
    const int ng = d.ng, nx = d.nx;
    int ix = p.ix, ig = p.ig;
    double x = p.x, mu = p.mu, e = p.e, e0 = p.e0, wl = p.wl, vx = p.vx;
    int ic = ix;
    double r1, r2, db, dcol, dthm, ddop, dd;
    int ixnext;
    erd = 0.0; jrd = 0.0; eamp = 0.0;
    double elabfact = 1.0 - mu * vx * CINV;
    const double elabfactold = elabfact;
    const double xl = XA(ix), xr = XA(ix + 1);
    const double far = 2.0 * fabs(sqrt(xr * xr - xl * xl));
    if (ix == 1 || mu >= -sqrt(1.0 - (xl / x) * (xl / x))) {
        db = fabs(sqrt(xr * xr - (1.0 - mu * mu) * x * x) - mu * x);
        ixnext = ix + 1;
    } else {
        db = fabs(sqrt(xl * xl - (1.0 - mu * mu) * x * x) + mu * x);
        ixnext = ix - 1;
    }
    if (db != db) return 1;
    const double sg = SIGT(ic), cp = CAPT(ig, ic);
    if (sg > 0.0) { r1 = s.draw(); dthm = p_path<M>(r1, elabfact * sg); }
    else dthm = far;
    if (cp <= 0.0) dcol = far;
    else { r1 = s.draw(); dcol = p_path<M>(r1, elabfact * cp); }
    const double dvdr = d.dvdr[ix - 1];
    const double vhelp = VXA(ix) - dvdr * xl;
    if (fabs(vx - vhelp) / PC_C >= 0.1) {
        // Newton-Raphson Doppler distance (transport11.f90:159-199)
        const double epstol = 1e-3;
        int it = 0;
        double shelp = db;
        double xhelp = sqrt(x * x + shelp * shelp + 2.0 * shelp * x * mu);
        double muhelp = (mu * x + shelp) / xhelp;
        double phi = dvdr * mu * x - PC_C * (1 - wl * WLI(ig + 1)) + vhelp * muhelp + dvdr * shelp;
        double dphi = dvdr + vhelp * (1 - muhelp * muhelp) / xhelp;
        ddop = shelp - phi / dphi;
        while ((fabs(ddop - shelp) > ddop * epstol) && it < 100) {
            it = it + 1;
            shelp = ddop;
            xhelp = sqrt(x * x + shelp * shelp + 2.0 * shelp * x * mu);
            muhelp = (mu * x + shelp) / xhelp;
            if (dvdr > 0 && ig < ng)
                phi = dvdr * mu * x - PC_C * (1 - wl * WLI(ig + 1)) + vhelp * muhelp + dvdr * shelp;
            else if (dvdr < 0 && ig > 1)
                phi = dvdr * mu * x - PC_C * (1 - wl * WLI(ig)) + vhelp * muhelp + dvdr * shelp;
            else { ddop = far; break; }
            dphi = dvdr + vhelp * (1 - muhelp * muhelp) / xhelp;
            ddop = shelp - phi / dphi;
        }
        if (ddop <= 0.0 || it == 100) ddop = far;
    } else {
        if (dvdr > 0 && ig < ng) ddop = d.c3divv[ix - 1] * (elabfact - wl * WLI(ig + 1));
        else if (dvdr < 0 && ig > 1) ddop = d.c3divv[ix - 1] * (elabfact - wl * WLI(ig));
        else ddop = far;
        if (ddop < 0.0) ddop = far;
    }
    // minloc over [dcol, dthm, ddop, db]
    int idist = 1;
    dd = dcol;
    if (dthm < dd) { dd = dthm; idist = 2; }
    if (ddop < dd) { dd = ddop; idist = 3; }
    if (db < dd) { dd = db; idist = 4; }
    p.idist = idist;
    if (dcol != dcol || dthm != dthm || ddop != ddop) return 3;
    if (dd < 0.0) return 4;
    const double xold = x;
    x = sqrt(x * x + dd * dd + 2.0 * dd * x * mu);
    vx = (VXA(ix) * (xr - x) + VXA(ix + 1) * (x - xl)) / DXX(ix);
    if (x == 0.0) mu = 1.0;
    else mu = (xold * mu + dd) / x;
    erd = e * elabfact * elabfact * dd * CINV;
    jrd = PC_C * erd;
    elabfact = 1.0 - mu * vx * CINV;
    int ret = 0;
    if (x < XA(1)) {
        x = XA(1);
        vx = VXA(1);
        ix = 1;
        r1 = s.draw();
        mu = 1.0 - r1;
        goto done;
    }
    if (dd == dthm || dd == dcol) {
        r1 = s.draw();
        if (dd == dthm && d.thm_aniso) mu = thomson_aniso_mu(r1);
        else mu = 1.0 - 2.0 * r1;
        mu = (mu + vx * CINV) / (1.0 + vx * mu * CINV);
    } else if (dd == db) {
        if (mu >= 0.0 && ix == nx) { p.stat = 1; goto done; }
    }
    if (dd == dthm) {
        wl = wl * (1.0 - mu * vx * CINV) / elabfact;
        const double help = elabfact / (1.0 - mu * vx * CINV);
        evelo = evelo + e * (1.0 - help);
        e = e * help;
        e0 = e0 * help;
    } else if (dd == dcol) {
        r1 = s.draw();                 // unused draw (transport11.f90:313)
        const double help = elabfact / (1.0 - mu * vx * CINV);
        evelo = evelo + e * (1.0 - help);
        e = e * help;
        e0 = e0 * help;
        if (ng != 1) {
            r1 = s.draw();
            int fl;
            ig = sample_entry(d, TAB_E, ic, 0, r1, bt_row(d, TAB_E, ic)[bt_key(r1)], fl);
            if (fl < 0) fl = (int)group_flags(d, ig, ic);
            if (!(fl & (int)BT_THIN)) {
                p.itype = 2;
                p.nflag = NF_VALID | ((uint32_t)fl & BT_LUMP);
            }
        }
        if (p.itype == 2) {
            evelo = evelo + e * vx * mu * CINV;
            e = e * (1.0 - vx * mu * CINV);
            e0 = e0 * (1.0 - vx * mu * CINV);
        } else {
            r1 = s.draw();
            wl = 1.0 / ((1.0 - r1) * WLI(ig) + r1 * WLI(ig + 1));
            wl = wl * (1.0 - vx * mu * CINV);
        }
    } else if (dd == db && ixnext > ix && ix == nx) {
        // outer boundary reached with a rounding-negative cosine: escapes (the reference
        // would index the non-existent cell nx+1 here)
        p.stat = 1;
        goto done;
    } else if (dd == db && ixnext > ix) {
        const int l = ix + 1;
        if ((SIGT(l) + CAPT(ig, l)) * DXX(ix + 1) < d.tauddmc || d.puretran) {
            x = XA(ix + 1);
            vx = VXA(ix + 1);
            ix = ix + 1;
        } else {
            mu = (mu - vx * CINV) / (1.0 - vx * mu * CINV);
            double help = (CAPT(ig, l) + SIGT(l)) * DXX(ix + 1);
            help = 4.0 / (3.0 * help + 6.0 * PC_DEXT);
            r1 = s.draw();
            if (r1 < help * (1.0 + 1.5 * fabs(mu))) {
                p.itype = 2;
                p.nflag = NF_VALID | lump_flag(d, CAPT(ig, l), DXX(ix + 1), l);
                evelo = evelo + e * (1.0 - elabfact);
                e = e * elabfact;
                e0 = e0 * elabfact;
                wl = wl / elabfact;
                ix = ix + 1;
            } else {
                r1 = s.draw();
                r2 = s.draw();
                mu = -(r1 > r2 ? r1 : r2);
                mu = (mu + vx * CINV) / (1.0 + vx * mu * CINV);
                x = XA(ix + 1);
                vx = VXA(ix + 1);
            }
        }
    } else if (dd == db) {
        const int l = ix - 1;
        if ((SIGT(l) + CAPT(ig, l)) * DXX(ix - 1) < d.tauddmc || d.puretran) {
            x = XA(ix);
            vx = VXA(ix);
            ix = ix - 1;
        } else {
            mu = (mu - vx * CINV) / (1.0 - vx * mu * CINV);
            if (d.ampfact && mu < 0.0) {
                // amplification factor (transport11.f90:410-421)
                double h = 1.0 / fabs(mu);
                h = h < 100.0 ? h : 100.0;
                evelo = evelo - e * 2.0 * (0.55 * h - 1.25 * fabs(mu)) * vx * CINV;
                eamp = e * 2.0 * 0.55 * (h - 2.0 > 0.0 ? h - 2.0 : 0.0) * vx * CINV;
                h = h < 2.0 ? h : 2.0;
                e0 = e0 * (1.0 + 2.0 * (0.55 * h - 1.25 * fabs(mu)) * vx * CINV);
                e = e * (1.0 + 2.0 * (0.55 * h - 1.25 * fabs(mu)) * vx * CINV);
            }
            double help = (CAPT(ig, l) + SIGT(l)) * DXX(ix - 1);
            help = 4.0 / (3.0 * help + 6.0 * PC_DEXT);
            r1 = s.draw();
            if (r1 < help * (1.0 + 1.5 * fabs(mu))) {
                p.itype = 2;
                p.nflag = NF_VALID | lump_flag(d, CAPT(ig, l), DXX(ix - 1), l);
                evelo = evelo + e * (1.0 - elabfact);
                e = e * elabfact;
                e0 = e0 * elabfact;
                wl = wl / elabfact;
                ix = ix - 1;
            } else {
                r1 = s.draw();
                r2 = s.draw();
                mu = r1 > r2 ? r1 : r2;
                mu = (mu + vx * CINV) / (1.0 + vx * mu * CINV);
                x = XA(ix);
                vx = VXA(ix);
            }
        }
    } else if (dd == ddop) {
        if (dvdr > 0) {
            if (ig < ng) {
                ig = ig + 1;
                wl = (WLG(ig) + 1e-6 * (WLG(ig + 1) - WLG(ig))) * elabfactold;
            } else {
                r1 = s.draw();
                wl = 1.0 / (r1 * WLI(ng + 1) + (1.0 - r1) * WLI(ng));
                wl = wl * elabfact;
            }
        } else if (dvdr < 0) {
            if (ig > 1) {
                wl = (WLG(ig) + 1e-6 * (WLG(ig - 1) - WLG(ig))) * elabfactold;
                ig = ig - 1;
            } else {
                r1 = s.draw();
                wl = 1.0 / (r1 * WLI(1) + (1.0 - r1) * WLI(2));
                wl = wl * elabfact;
            }
        }
        if (((SIGT(ic) + CAPT(ig, ic)) * DXX(ix) >= d.tauddmc) && !d.puretran) {
            p.itype = 2;
            p.nflag = NF_VALID | lump_flag(d, CAPT(ig, ic), DXX(ix), ic);
            evelo = evelo + e * vx * mu * CINV;
            e = e * elabfact;
            e0 = e0 * elabfact;
            wl = wl / elabfact;
        }
    } else {
        ret = 17;
    }
done:
    p.x = x; p.mu = mu; p.e = e; p.e0 = e0; p.wl = wl; p.vx = vx;
    p.ix = ix; p.ig = ig;
    return ret;
}

// resolved leakage opacity through the left (side<0) / right face (diffusion11.f90:204-239)
__device__ __forceinline__ double resleak_t(const Dev &d, int ix, int g, int side)
{
    const int ng = d.ng, ic = ix;
    if (side < 0) {
        const int l = ix - 1;
        if ((CAPT(g, l) + SIGT(l)) * DXX(ix - 1) < d.tauddmc) {
            double mfp = (CAPT(g, ic) + SIGT(ic)) * DXX(ix);
            double ppl = 4.0 / (3.0 * mfp + 6.0 * PC_DEXT);
            return 1.5 * ppl * (XA(ix) * XA(ix)) / DX3X(ix);
        }
        double mfp = (SIGT(ic) + CAPT(g, ic)) * DXX(ix) + (SIGT(l) + CAPT(g, l)) * DXX(ix - 1);
        return 2.0 * (XA(ix) * XA(ix)) / (mfp * DX3X(ix));
    }
    if (ix == d.nx) {
        double mfp = (CAPT(g, ic) + SIGT(ic)) * DXX(ix);
        double ppr = 4.0 / (3.0 * mfp + 6.0 * PC_DEXT);
        return 1.5 * ppr * (XA(ix + 1) * XA(ix + 1)) / DX3X(ix);
    }
    const int l = ix + 1;
    if ((CAPT(g, l) + SIGT(l)) * DXX(ix + 1) < d.tauddmc) {
        double mfp = (CAPT(g, ic) + SIGT(ic)) * DXX(ix);
        double ppr = 4.0 / (3.0 * mfp + 6.0 * PC_DEXT);
        return 1.5 * ppr * (XA(ix + 1) * XA(ix + 1)) / DX3X(ix);
    }
    double mfp = (SIGT(ic) + CAPT(g, ic)) * DXX(ix) + (SIGT(l) + CAPT(g, l)) * DXX(ix + 1);
    return 2.0 * (XA(ix + 1) * XA(ix + 1)) / (mfp * DX3X(ix));
}

// uniform position in the shell with the single-precision 1/3 exponent, clamped (:336-339)
template <bool M>
__device__ __forceinline__ double shell_position(const Dev &d, int ix, double r1)
{
    const double xl = XA(ix), xr = XA(ix + 1);
    double x = p_shell_root<M>(r1, xl, xr);
    x = x > xr ? xr : x;
    x = x < xl ? xl : x;
    return x;
}

// diffusion11.f90:243-283: the denominator and event thresholds of one DDMC step
__device__ __forceinline__ DdmcProb ddmc_prob(double opl, double opr, double emitlump,
                                              double caplump, double doplump)
{
    double denom = opl + opr + (1.0 - emitlump) * caplump + doplump;
    denom = 1.0 / denom;
    const double pl = opl * denom, pr = opr * denom, pdop = doplump * denom;
    DdmcProb q;
    q.denom = denom;
    q.t1 = 0.0 + pdop;
    q.t2 = 0.0 + pdop + pl;
    q.t3 = 0.0 + pdop + (pl + pr);
    return q;
}

// The per-event constants of diffusion11.f90:95-152 and 204-239 depend only on the cell
// and, for an unlumped group, on the group; they are evaluated once per iteration here,
// one thread per (group, cell), and the step loads them.
__global__ __launch_bounds__(256) void k_ddmc_probs(Dev d)
{
    const int ng = d.ng;
    const int ig = blockIdx.x * 256 + threadIdx.x + 1, ic = blockIdx.y + 1, ix = ic;
    if (ig > ng) return;
    const CellDDMC cc = d.cell[ic - 1];
    const double cp = CAPT(ig, ic);
    // diffusion11.f90:175-182: grd_cap(ig)*capgreyinv in LTE, grd_capemit(ig)*capemitgreyinv in NLTE
    const double ce = (double)d.capem32[(size_t)(ic - 1) * ng + ig - 1];
    const double emitlump = specint0(TINV(ic), d.wlinv, ig) * cc.emgreyinv * ce;
    const double sa = d.specarr[(size_t)(ic - 1) * ng + ig - 1];
    double doplump;
    if (ig != ng && DIVV(ix) >= 0)
        doplump = dopspeccalc(TINV(ic), d.wlinv, ig) * DIVV(ix) / (sa * 3 * PC_C);
    else if (ig != 1 && DIVV(ix) < 0)
        doplump = -1 * dopspeccalc(TINV(ic), d.wlinv, ig - 1) * DIVV(ix) / (sa * 3 * PC_C);
    else
        doplump = 0.0;
    const double opl = (ix == 1) ? 0.0 : resleak_t(d, ix, ig, -1);
    const double opr = resleak_t(d, ix, ig, +1);
    d.gprob[(size_t)(ic - 1) * ng + ig - 1] = ddmc_prob(opl, opr, emitlump, cp, doplump);
    if (ig == 1) d.cprob[ic - 1] = ddmc_prob(OPL(1, ic), OPL(2, ic), cc.emitlump, cc.caplump, cc.doplump);
}

void launch_ddmc_probs(slgpu_ctx *c)
{
    hipLaunchKernelGGL(k_ddmc_probs, dim3((c->ng + 255) / 256, c->nx), dim3(256), 0, c->stream,
                       c->d);
}

// TRANSPORT1/diffusion11.f90:91-631 — one DDMC event.  Every cell constant is evaluated for
// the current cell.  The four events (Doppler, left leak, right leak, effective scatter,
// :283-620) run through ONE code path: the event selects the sampling table, the
// destination cell and the IMC-conversion rule by data, while the random draws keep the
// reference's order per event.  That keeps the lanes of a wave on the same instructions.
enum { EV_DOP = 0, EV_LEFT = 1, EV_RIGHT = 2, EV_SCAT = 3 };

template <bool M>
__device__ int diffusion11(const Dev &d, Ptcl &p, StepRng &s, double &erd, double &jrd,
                           double &evelo, double &pend_f, double &pend_r2)
{
    const int ng = d.ng, nx = d.nx;
    int ix = p.ix, ig = p.ig;
    const int ic = ix;
    double x = p.x, mu = p.mu, e = p.e, e0 = p.e0, wl = p.wl, vx = p.vx;
    double r1;
    erd = 0.0; jrd = 0.0;
    const double dist = DXX(ix);
    const int nlump = d.cell[ic - 1].nlump;
    // The step's first three draws.  Draw 2 samples the group of every effective scatter
    // (and of lumped events), so with the flags the previous step left, the thresholds and
    // that draw's bucket entry are requested together: one memory round trip per step
    // instead of opacity -> thresholds -> entry.  Draws are counted on (step, index), so
    // drawing early does not change any value.
    const double rt = s.draw(), rev = s.draw(), r2 = s.draw();
    const bool known = (p.nflag & NF_VALID) != 0;
    uint32_t lump;                                             // BT_LUMP of (ig, ic)
    if (known) {
        lump = p.nflag & BT_LUMP;
    } else {
        const double cp = CAPT(ig, ic);
        if ((SIGT(ic) + cp) * dist < d.tauddmc) return 100;
        lump = (cp * dist >= d.taulump && nlump > 0) ? BT_LUMP : 0u;
    }
    const int glump = lump ? nlump : 0;                        // :156-160
    // the cell's lumped thresholds or the group's own
    const DdmcProb q = lump ? d.cprob[ic - 1] : d.gprob[(size_t)(ic - 1) * ng + (ig - 1)];
    const int tscat = lump ? TAB_S : TAB_E;
    const uint32_t escat = known ? bt_row(d, tscat, ic)[bt_key(r2)] : 0u;
    const double denom = q.denom;
    const double ddmct = fabs(p_log<M>(rt) * denom * CINV);
    erd = e * ddmct;
    jrd = PC_C * erd;
    int ev;
    if (rev >= 0.0 && rev < q.t1) ev = EV_DOP;
    else if (rev >= q.t1 && rev < q.t2) ev = EV_LEFT;
    else if (rev >= q.t2 && rev < q.t3) ev = EV_RIGHT;
    else ev = EV_SCAT;
    // keep the entry's load ahead of the branches that use it
    asm volatile("" ::"v"(escat));
    p.idist = ev == EV_LEFT ? -3 : (ev == EV_RIGHT ? -4 : (ev == EV_SCAT ? -2 : p.idist));
    // draws after the first three continue the step's sequence; draw 2 is taken by the
    // first consumer that needs a fourth number in the reference's order
    bool r2_free = true;
    auto next = [&]() -> double {
        if (r2_free) { r2_free = false; return r2; }
        return s.draw();
    };
    int ret = 0;
    bool adiabatic = true;
    uint32_t nflag = NF_VALID | lump;                          // particle stays in (ig, ic)
    if (ev == EV_LEFT && ix == 1) {
        adiabatic = false;                                     // :362-365 returns
    } else if (ev == EV_SCAT && glump == ng) {
        ret = 102;                                             // :552-556
        adiabatic = false;
    } else {
        const bool flux = (ev == EV_RIGHT && ix == nx);        // :438-478
        if (flux) {
            p.stat = 1;
            const double ra = next(), rb = next();
            mu = ra > rb ? ra : rb;
        }
        // group sampling: lumped particles draw from the cell's table of this event,
        // unlumped effective scatters from the emission probabilities (emitgroup)
        int iiig = ig, fl = -1;
        if (glump > 0 || ev == EV_SCAT) {
            const bool pre = known && r2_free && ev == EV_SCAT;
            r1 = next();
            const int t = glump == 0 ? TAB_E
                        : (ev == EV_LEFT ? TAB_L : (ev == EV_RIGHT ? TAB_R : (ev == EV_DOP ? TAB_D : TAB_S)));
            const uint32_t ent = pre ? escat : bt_row(d, t, ic)[bt_key(r1)];
            iiig = sample_entry(d, t, ic, nlump, r1, ent, fl);
        }
        if (flux) {
            ig = iiig;
            r1 = next();
            wl = 1.0 / (r1 * WLI(ig + 1) + (1.0 - r1) * WLI(ig));
            x = XA(nx + 1);
            vx = VXA(nx + 1);
            mu = (mu + vx * CINV) / (1.0 + vx * mu * CINV);
            mu = mu > 1.0 ? 1.0 : mu;
            const double help = 1.0 / (1.0 - mu * vx * CINV);
            evelo = evelo + e * (1.0 - help);
            wl = wl / help;
            e = e * help;
            e0 = e0 * help;
            adiabatic = false;
        } else {
            // destination group and the cell whose optical depth decides IMC conversion
            int ig2 = iiig, lc = ic;
            if (ev == EV_DOP) {
                if (DIVV(ix) >= 0) { ig2 = iiig + 1; wl = WLG(ig2); ig2 = ig2 > ng ? ng : ig2; }
                else { ig2 = iiig - 1; wl = WLG(ig2 + 1); ig2 = ig2 < 1 ? 1 : ig2; }
            } else if (ev == EV_LEFT) {
                lc = ix - 1;
            } else if (ev == EV_RIGHT) {
                lc = ix + 1;
            }
            if (fl < 0) fl = (int)group_flags(d, ig2, lc);
            nflag = NF_VALID | ((uint32_t)fl & BT_LUMP);
            if (fl & (int)BT_THIN) {
                // conversion to IMC: recorded here, sampled by imc_entry at the next pass
                p.itype = 1;
                const uint32_t kind = ev == EV_SCAT ? PEND_SCAT
                                    : (ev == EV_LEFT ? PEND_LEFT : (ev == EV_RIGHT ? PEND_RIGHT : PEND_DOP));
                nflag = (kind << 1) | (r2_free ? NF_R2FREE : 0u) | ((uint32_t)fl & BT_LUMP);
                if (r2_free) pend_r2 = r2;
            }
            ig = ig2;
            if (ev == EV_LEFT) ix = ix - 1;
            else if (ev == EV_RIGHT) ix = ix + 1;
        }
    }
    p.nflag = nflag;
    if (adiabatic) {
        if (nflag & NF_PEND) {
            // after the conversion's frame change, at the next pass (its position lies in
            // the shell, so the inner-boundary test below never fires for it)
            pend_f = p_adiabatic<M>(DIVV(ix), ddmct);
        } else if (x < XA(1)) {
            p.stat = 2;                                        // :623-626
        } else {
            const double f = p_adiabatic<M>(DIVV(ix), ddmct);
            evelo = evelo + e * (1.0 - f);
            e = e * f;
            e0 = e0 * f;
        }
    }
    p.x = x; p.mu = mu; p.e = e; p.e0 = e0; p.wl = wl; p.vx = vx;
    p.ix = ix; p.ig = ig;
    return ret;
}

// The recorded DDMC -> IMC conversion of step `step` (diffusion11.f90:329-355, 402-423,
// 515-538, 587-618): the new wavelength, direction and position, the frame change, then the
// step's adiabatic factor f, in the reference's order.  Its draws are that step's: draw 2 when
// it was still unused (r2), then draws 3.. (block 1 of the step).
template <bool M>
__device__ void imc_entry(const Dev &d, Ptcl &p, const StepRng &s, uint32_t step, double &evelo,
                          double f, double r2)
{
    const uint32_t kind = (p.nflag & NF_PEND) >> 1;
    bool r2_free = (p.nflag & NF_R2FREE) != 0;
    uint32_t o[4];
    Philox::block(s.pid, step, StepRng::PHASE, 1u, s.k0, s.k1, o);
    uint32_t j = 0;
    auto next = [&]() -> double {
        if (r2_free) { r2_free = false; return r2; }
        const double u = Philox::u42(j == 0 ? o[0] : (j == 1 ? o[1] : o[2]), (o[3] >> (10u * j)) & 0x3FFu);
        ++j;
        return u;
    };
    const int ix = p.ix, ig = p.ig;
    double x, mu, vx, e = p.e, e0 = p.e0, wl = p.wl;
    const bool face = kind == PEND_LEFT || kind == PEND_RIGHT;
    if (kind == PEND_SCAT) {
        const double r1 = next();
        wl = 1.0 / ((1.0 - r1) * WLI(ig) + r1 * WLI(ig + 1));
    } else if (face) {
        const double r1 = next();
        wl = 1.0 / (r1 * WLI(ig + 1) + (1.0 - r1) * WLI(ig));
    }
    const double ra = next(), rb = next();
    if (face) {
        const int xi = kind == PEND_LEFT ? ix + 1 : ix;      // the face crossed into cell ix
        x = XA(xi);
        vx = VXA(xi);
        mu = ra > rb ? ra : rb;
        if (kind == PEND_LEFT) mu = -mu;
    } else {
        mu = 1.0 - 2.0 * ra;
        x = shell_position<M>(d, ix, rb);
        vx = (VXA(ix) * (XA(ix + 1) - x) + VXA(ix + 1) * (x - XA(ix))) / DXX(ix);
    }
    mu = (mu + vx * CINV) / (1.0 + vx * mu * CINV);
    const double help = 1.0 / (1.0 - vx * mu * CINV);
    evelo = evelo + e * (1.0 - help);
    e = e * help;
    e0 = e0 * help;
    wl = wl * (1.0 - vx * mu * CINV);
    evelo = evelo + e * (1.0 - f);
    e = e * f;
    e0 = e0 * f;
    p.x = x; p.mu = mu; p.vx = vx; p.e = e; p.e0 = e0; p.wl = wl;
    p.nflag = NF_VALID | (p.nflag & BT_LUMP);
}

// Block totals in LDS, indexed like Dev::tot and Dev::ct.  Whatever changes once per
// particle (its start, its end, the rare warnings and amplification energies) is added
// here at once; only the per-step sums stay in registers (evelo, nddmc, nswap), which
// keeps the persistent lane state small enough to avoid spilling.  The IMC step count is
// the particles' total steps minus the DDMC steps.
struct BlockTotals {
    double tot[7];                 // eout, evelo, -, estart, edead, esrc, eamp
    unsigned long long ct[9];      // nact, total steps (then nimc), nddmc, nmax, nswap,
                                   // nflux, ndead, nerr, nwarn
};

__device__ __forceinline__ void flush_totals(const Dev &d, BlockTotals &b, double evelo,
                                             uint32_t nddmc, uint32_t nswap)
{
    const double ev = wave_sum(evelo);
    unsigned long long nd = nddmc, ns = nswap;
    for (int off = 32; off > 0; off >>= 1) {
        nd += __shfl_xor(nd, off, 64);
        ns += __shfl_xor(ns, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (ev != 0.0) atomicAdd(&b.tot[1], ev);
        atomicAdd(&b.ct[2], nd);
        atomicAdd(&b.ct[4], ns);
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        const double v = b.tot[threadIdx.x];
        if (v != 0.0) atomicAdd(&d.tot[threadIdx.x], v);
    } else if (threadIdx.x >= 64 && threadIdx.x < 64 + 9) {
        const int q = threadIdx.x - 64;
        unsigned long long v = b.ct[q];
        if (q == 1) v -= b.ct[2];                                // IMC steps
        if (q == 3) atomicMax(&d.ct[3], v);
        else if (v) atomicAdd(&d.ct[q], v);
    }
}

// particle_advance.f90:140-294 as a persistent kernel: each lane takes the next particle
// from a global counter, runs it event by event to escape or death (145-202 on start,
// 271-292 at the end), then takes another, so lanes stay busy until the pool is empty.
// Lanes that leave the loop drop out of the votes; a pass always runs the type of at
// least one remaining lane, so every particle advances.
// Occupancy target per precision (waves per SIMD), measured on c2: the precision form runs
// best at 4 (VGPRs capped at 128), the FP64 form at 3 (capped at 168); the compiler's own
// choice (2 and 3) is 6-11% slower.
// A wave runs an IMC pass when n_imc * SLGPU_VOTE_IMC >= n_ddmc * SLGPU_VOTE_DDMC.  IMC
// lanes wait out DDMC passes, so letting them go once they are a third of the live lanes
// (2:1) beat a plain majority by 3-4% on c2 (1:2 and 4:1 were slower).
#ifndef SLGPU_VOTE_IMC
#define SLGPU_VOTE_IMC 2
#endif
#ifndef SLGPU_VOTE_DDMC
#define SLGPU_VOTE_DDMC 1
#endif
#ifndef SLGPU_WAVES_MIXED
#define SLGPU_WAVES_MIXED 4
#endif
#ifndef SLGPU_WAVES_FP64
#define SLGPU_WAVES_FP64 3
#endif
template <bool M>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(M ? SLGPU_WAVES_MIXED : SLGPU_WAVES_FP64, 8)))
void k_transport(Dev d, uint32_t iteration)
{
    extern __shared__ double s_tally[];
    const int ng = d.ng, nx = d.nx;
    __shared__ BlockTotals bt;
    __shared__ double s_pend_f[256], s_pend_r2[256];       // a lane's recorded conversion
    for (int k = threadIdx.x; k < nx; k += blockDim.x) s_tally[k] = 0.0;
    if (threadIdx.x < 7) bt.tot[threadIdx.x] = 0.0;
    if (threadIdx.x < 9) bt.ct[threadIdx.x] = 0;
    __syncthreads();
    double evelo = 0.0;
    uint32_t nddmc = 0, nswap = 0;
    int run_ic = -1;
    double run_e = 0.0;
    double *jrad = d.jrad + (size_t)(blockIdx.x % d.jrad_copies) * nx * ng;
    Ptcl p;
    StepRng s;
    s.key(d.seed, iteration);
    int64_t ip = -1, istep = 0;
    bool have = false;
    for (;;) {
        if (!have) {
            ip = (int64_t)atomicAdd(d.queue, 1ull);
            if (ip >= d.npart) break;
            p.x = d.px[ip]; p.mu = d.pmu[ip]; p.e = d.pe[ip]; p.e0 = d.pe0[ip]; p.wl = d.pwl[ip];
            atomicAdd(&bt.tot[5], p.e);
            atomicAdd(&bt.ct[0], 1ull);
            p.ix = binsrch(p.x, d.xarr, nx + 1, false);
            p.ig = binsrch(p.wl, d.wl, ng + 1, false);
            if (p.ix < 1 || p.ig < 1) {                        // NaN position or wavelength
                atomicAdd(&bt.ct[7], 1ull); atomicAdd(&bt.ct[6], 1ull);
                atomicAdd(&bt.tot[4], p.e);
                d.pstat[ip] = 2; d.perr[ip] = 1; d.pnstep[ip] = 0;
                continue;
            }
            const double tau = (SIGT(p.ix) + CAPT(p.ig, p.ix)) * DXX(p.ix);
            p.itype = (d.puretran || tau < d.tauddmc) ? 1 : 2;
            const double h1 = XA(p.ix), h2 = XA(p.ix + 1);
            p.vx = (VXA(p.ix) * (h2 - p.x) + VXA(p.ix + 1) * (p.x - h1)) / (h2 - h1);
            if (p.itype == 1) {
                const double labfact = 1.0 - p.vx * p.mu / PC_C;
                p.wl = p.wl * labfact;
                p.e = p.e / labfact;
                p.e0 = p.e0 / labfact;
            }
            atomicAdd(&bt.tot[3], p.e);
            p.stat = 0;
            p.idist = 0;
            p.nflag = 0;
            s.init((uint32_t)ip);
            istep = 0;
            have = true;
        }
        // One event.  The wave runs one event type per pass, chosen by the vote above the
        // kernel; lanes of the other type wait a pass, so each pass executes one body.
        const bool imc = p.itype == 1 || d.puretran;
        const int n_imc = __popcll(__ballot(imc)), n_ddmc = __popcll(__ballot(!imc));
        if (imc != (n_imc * SLGPU_VOTE_IMC >= n_ddmc * SLGPU_VOTE_DDMC)) continue;
        istep++;
        if (imc && (p.nflag & NF_PEND))
            imc_entry<M>(d, p, s, (uint32_t)(istep - 1), evelo, s_pend_f[threadIdx.x], s_pend_r2[threadIdx.x]);
        s.begin_step((uint32_t)istep);
        const int icold = p.ix, igold = p.ig;
        double erd, jrd;
        int ierr;
        if (imc) {
            double ea;
            ierr = transport11<M>(d, p, s, erd, jrd, ea, evelo);
            if (ea != 0.0) {                                   // particle_advance.f90:216-217
                atomicAdd(&d.eamp[icold - 1], ea);
                atomicAdd(&bt.tot[6], ea);
            }
            if (p.itype != 1) nswap++;
        } else {
            nddmc++;
            ierr = diffusion11<M>(d, p, s, erd, jrd, evelo, s_pend_f[threadIdx.x], s_pend_r2[threadIdx.x]);
            if (p.itype == 1) nswap++;
        }
        // grd_jrad: one FP64 atomic per event into the block's replica (the group changes on
        // almost every step, so merging a particle's events would save no atomics); the
        // energy tally merges a particle's consecutive events in one cell before its LDS atomic
        if (jrd != 0.0) atomicAdd(&jrad[(size_t)(icold - 1) * ng + igold - 1], jrd);
        if (icold != run_ic) {
            if (run_ic > 0) atomicAdd(&s_tally[run_ic - 1], run_e);
            run_ic = icold; run_e = 0.0;
        }
        run_e += erd;
        if (p.itype == 1 && p.stat == 0 && !(p.nflag & NF_PEND)) {   // particle_advance.f90:234-250
            if (p.x > XA(p.ix + 1) || p.x < XA(p.ix) || p.x != p.x) {
                if (ierr == 0) ierr = -99;
            }
        }
        if (ierr < 0) atomicAdd(&bt.ct[8], 1ull);
        if (ierr == 0 && p.stat == 0 && istep >= d.max_steps) ierr = 200;
        if (ierr > 0) { atomicAdd(&bt.ct[7], 1ull); p.stat = 2; }
        if (p.stat != 0) {
            if (p.nflag & NF_PEND)                             // stopped by max_steps right after
                imc_entry<M>(d, p, s, (uint32_t)istep, evelo, s_pend_f[threadIdx.x], s_pend_r2[threadIdx.x]);
            atomicMax(&bt.ct[3], (unsigned long long)istep);
            atomicAdd(&bt.ct[1], (unsigned long long)istep);
            if (p.stat == 2) { atomicAdd(&bt.ct[6], 1ull); atomicAdd(&bt.tot[4], p.e); }
            else { atomicAdd(&bt.ct[5], 1ull); atomicAdd(&bt.tot[0], p.e); }
            d.pstat[ip] = (uint8_t)p.stat;
            d.px[ip] = p.x; d.pmu[ip] = p.mu; d.pwl[ip] = p.wl; d.pe[ip] = p.e; d.pe0[ip] = p.e0;
            d.pig[ip] = p.ig; d.pix[ip] = p.ix;
            d.pnstep[ip] = istep;
            d.perr[ip] = ierr > 0 ? ierr : 0;
            have = false;
        }
    }
    if (run_ic > 0) atomicAdd(&s_tally[run_ic - 1], run_e);
    __syncthreads();
    for (int k = threadIdx.x; k < nx; k += blockDim.x)
        if (s_tally[k] != 0.0) atomicAdd(&d.tally[k], s_tally[k]);
    flush_totals(d, bt, evelo, nddmc, nswap);
}

// fluxtally.f:30-54 (flux grid == group grid; polar/azimuth bins are single)
__global__ __launch_bounds__(256) void k_fluxtally(Dev d)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.npart) return;
    if (d.pstat[p] != 1) return;
    const int ig = binsrch(d.pwl[p], d.wl, d.ng + 1, false);
    if (ig < 1) return;                  // NaN wavelength: not binned
    const double e = d.pe[p];
    atomicAdd(&d.lum[ig - 1], e);
    atomicAdd(&d.dev[ig - 1], e * e);
    atomicAdd(&d.num[ig - 1], 1);
}

void launch_sources(slgpu_ctx *c, uint32_t iteration)
{
    int64_t n = c->d.npart;
    if (n == 0) return;
    hipLaunchKernelGGL(k_sources, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       c->d, iteration);
}

void launch_transport(slgpu_ctx *c, uint32_t iteration)
{
    int dev = 0, ncu = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    int64_t blocks = (int64_t)ncu * 8;
    const int64_t need = (c->d.npart + 255) / 256;
    if (need < blocks) blocks = need > 0 ? need : 1;
    if (c->d.precision == SLGPU_MIXED)
        hipLaunchKernelGGL(k_transport<true>, dim3((unsigned)blocks), dim3(256),
                           c->nx * sizeof(double), c->stream, c->d, iteration);
    else
        hipLaunchKernelGGL(k_transport<false>, dim3((unsigned)blocks), dim3(256),
                           c->nx * sizeof(double), c->stream, c->d, iteration);
}

// the transport's group samplers at arbitrary (r, cell), table path or direct search
__global__ __launch_bounds__(256) void k_sample_groups(Dev d, int t, int direct, int64_t n,
                                                       const double *r, const int *cell, int *out)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int ic = cell[k];
    if (ic < 1 || ic > d.nx) { out[k] = -1; return; }
    const int nlump = d.cell[ic - 1].nlump;
    if (t != TAB_E) {
        const LumpRow L = lump_row(d, t, ic, nlump);
        if (L.cnt == 0) { out[k] = 0; return; }
        if (direct) { out[k] = lump_group(d, L, ic, first_above(L.cum, 0, L.cnt, r[k])); return; }
    } else if (direct) {
        out[k] = emitgroup_direct(d, r[k], ic);
        return;
    }
    int thin;
    out[k] = sample_group(d, t, ic, nlump, r[k], thin);
}

void launch_sample_groups(slgpu_ctx *c, int t, int direct, int64_t n, const double *r,
                          const int *cell, int *out)
{
    hipLaunchKernelGGL(k_sample_groups, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       c->stream, c->d, t, direct, n, r, cell, out);
}

// fold the grd_jrad replicas into replica 0 (fixed order per element)
__global__ __launch_bounds__(256) void k_jrad_reduce(Dev d)
{
    const size_t n = (size_t)d.nx * d.ng;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (size_t)gridDim.x * blockDim.x) {
        double s = d.jrad[k];
        for (int r = 1; r < d.jrad_copies; ++r) s += d.jrad[(size_t)r * n + k];
        d.jrad[k] = s;
    }
}

void launch_jrad_reduce(slgpu_ctx *c)
{
    if (c->d.jrad_copies <= 1) return;
    const size_t n = (size_t)c->nx * c->ng;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_jrad_reduce, dim3((unsigned)blocks), dim3(256), 0, c->stream, c->d);
}

void launch_fluxtally(slgpu_ctx *c)
{
    int64_t n = c->d.npart;
    if (n == 0) return;
    hipLaunchKernelGGL(k_fluxtally, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       c->d);
}

}  // namespace slgpu
