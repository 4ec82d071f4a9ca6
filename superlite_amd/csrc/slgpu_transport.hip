// Source sampling, 1D-spherical IMC/DDMC particle transport and the flux tally on gfx950.
//
//   k_sources    SOURCE/interior_source.f90:57-168 + source_transformdirection.f90:23-46
//   k_transport  particle_advance.f90:140-294 with TRANSPORT1/transport11.f90:86-489
//                (one IMC event) and TRANSPORT1/diffusion11.f90:91-631 (one DDMC event)
//   k_fluxtally  fluxtally.f:30-54
//
// The transport kernel is persistent: every lane owns one particle at a time and takes
// one event per loop trip; a lane whose particle escaped or died fetches the next particle
// index from a global work counter, so the SIMD stays full while particles with very
// different event counts drain.  Every lane leaves the loop once the counter passes the
// particle count and its own particle is finished (each particle is capped at max_steps).
// Per-particle randomness is a Philox-4x32-10 stream keyed on (seed, iteration) and
// counted on (particle, draw, phase), so trajectories do not depend on the schedule.
//
// Tallies: the per-cell energy tally lives in LDS (nc <= 4096 doubles) per block and is
// flushed with FP64 atomics at the end; consecutive events of one particle in the same
// (cell, group) are merged in registers before the FP64 atomic into grd_jrad.
#include "slgpu_device.h"
#include "slgpu_internal.h"

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

// first index j in [lo, hi) with cum[j] > r (upper bound); hi if none
__device__ __forceinline__ int first_above(const double *__restrict__ cum, int lo, int hi, double r)
{
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cum[mid] > r) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// GRID/emitgroup.f:3-54 via the per-cell in-bin cumulative table
__device__ __forceinline__ int emitgroup(const Dev &d, double r, int ic)
{
    const int ng = d.ng;
    const double *ep = d.emitprob + (size_t)(ic - 1) * d.nep;
    const double r1 = r * d.capgrey[ic - 1];
    int iep = binsrch(r1, ep, d.nep, true);
    if (iep > d.nep - 1) iep = d.nep - 1;
    if (iep < 0) iep = 0;
    const int ig = iep * d.nepg + 1;
    const int igp1 = ig + d.nepg - 1 < ng ? ig + d.nepg - 1 : ng;
    // first group g in [ig, igp1-1] whose running sum exceeds r1, else igp1
    const double *cum = d.emitcum + (size_t)(ic - 1) * ng;
    int j = first_above(cum, ig - 1, igp1 - 1, r1);
    return j + 1;
}

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
};

// TRANSPORT1/transport11.f90:86-489 (trn_noampfact = .true.)
__device__ int transport11(const Dev &d, Ptcl &p, Philox &s, double &erd, double &jrd,
                           double &evelo)
{
    const int ng = d.ng, nx = d.nx;
    int ix = p.ix, ig = p.ig;
    double x = p.x, mu = p.mu, e = p.e, e0 = p.e0, wl = p.wl, vx = p.vx;
    int ic = ix;
    double r1, r2, db, dcol, dthm, ddop, dd;
    int ixnext;
    erd = 0.0; jrd = 0.0;
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
    if (sg > 0.0) { r1 = s.draw(); dthm = -log(r1) / (elabfact * sg); }
    else dthm = far;
    if (cp <= 0.0) dcol = far;
    else { r1 = s.draw(); dcol = -log(r1) / (elabfact * cp); }
    const double dvdr = (VXA(ix + 1) - VXA(ix)) / DXX(ix);
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
        if (dvdr > 0 && ig < ng) ddop = 3.0 * PC_C / DIVV(ix) * (elabfact - wl * WLI(ig + 1));
        else if (dvdr < 0 && ig > 1) ddop = 3.0 * PC_C / DIVV(ix) * (elabfact - wl * WLI(ig));
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
        mu = 1.0 - 2.0 * r1;
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
            ig = emitgroup(d, r1, ic);
            if (((SIGT(ic) + CAPT(ig, ic)) * DXX(ix) >= d.tauddmc) && !d.puretran) p.itype = 2;
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
            double help = (CAPT(ig, l) + SIGT(l)) * DXX(ix - 1);
            help = 4.0 / (3.0 * help + 6.0 * PC_DEXT);
            r1 = s.draw();
            if (r1 < help * (1.0 + 1.5 * fabs(mu))) {
                p.itype = 2;
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
__device__ __forceinline__ double shell_position(const Dev &d, int ix, double r1)
{
    const double third = (double)(1.0f / 3.0f);
    const double xl = XA(ix), xr = XA(ix + 1);
    double x = pow(r1 * (xr * xr * xr) + (1.0 - r1) * (xl * xl * xl), third);
    x = x > xr ? xr : x;
    x = x < xl ? xl : x;
    return x;
}

// TRANSPORT1/diffusion11.f90:91-631, every cell constant evaluated for the current cell
__device__ int diffusion11(const Dev &d, Ptcl &p, Philox &s, double &erd, double &jrd,
                           double &evelo)
{
    const int ng = d.ng, nx = d.nx;
    int ix = p.ix, ig = p.ig;
    const int ic = ix;
    double x = p.x, mu = p.mu, e = p.e, e0 = p.e0, wl = p.wl, vx = p.vx;
    double r1, r2;
    int iiig = ig;
    erd = 0.0; jrd = 0.0;
    const double dist = DXX(ix);
    const CellDDMC cc = d.cell[ic - 1];
    const double cp = CAPT(ig, ic);
    const int glump = (cp * dist >= d.taulump) ? cc.nlump : 0;
    if ((SIGT(ic) + cp) * dist < d.tauddmc) return 100;
    double emitlump, caplump, doplump;
    double opl, opr;
    if (glump > 0) {
        emitlump = cc.emitlump;
        caplump = cc.caplump;
        doplump = cc.doplump;
        opl = OPL(1, ic);
        opr = OPL(2, ic);
    } else {
        emitlump = specint0(TINV(ic), d.wlinv, ig) * cc.capgreyinv * cp;
        caplump = cp;
        const double sa = d.specarr[(size_t)(ic - 1) * ng + ig - 1];
        if (ig != ng && DIVV(ix) >= 0)
            doplump = dopspeccalc(TINV(ic), d.wlinv, ig) * DIVV(ix) / (sa * 3 * PC_C);
        else if (ig != 1 && DIVV(ix) < 0)
            doplump = -1 * dopspeccalc(TINV(ic), d.wlinv, ig - 1) * DIVV(ix) / (sa * 3 * PC_C);
        else
            doplump = 0.0;
        opl = (ix == 1) ? 0.0 : resleak_t(d, ix, ig, -1);
        opr = resleak_t(d, ix, ig, +1);
    }
    double denom = opl + opr + (1.0 - emitlump) * caplump + doplump;
    denom = 1.0 / denom;
    r1 = s.draw();
    const double ddmct = fabs(log(r1) * denom * CINV);
    erd = e * ddmct;
    jrd = PC_C * erd;
    r1 = s.draw();
    const double pl = opl * denom, pr = opr * denom;
    const double pdop = doplump * denom;
    const size_t row = (size_t)(ic - 1) * ng;
    int ret = 0;

    if (r1 >= 0.0 && r1 < 0.0 + pdop) {
        // Doppler shift out of the group / lump
        if (glump == 0) iiig = ig;
        else {
            r1 = s.draw();
            const int n = cc.nlump;
            int j = guided_first_above(d.cumD + row, n, d.guide + ((size_t)2 * nx + ic - 1) * GUIDE_K, d.guide_m[2 * nx + ic - 1], r1);
            if (DIVV(ix) >= 0) iiig = d.gorder[row + (j < n ? j : n - 1)];
            else iiig = d.gorder[row + (j < n ? n - 1 - j : 0)];
        }
        if (DIVV(ix) >= 0) {
            ig = iiig + 1;
            wl = WLG(ig);
            if (ig > ng) ig = ng;
        } else {
            ig = iiig - 1;
            wl = WLG(ig + 1);
            if (ig < 1) ig = 1;
        }
        if ((SIGT(ic) + CAPT(ig, ic)) * dist < d.tauddmc) {
            p.itype = 1;
            r1 = s.draw();
            mu = 1.0 - 2.0 * r1;
            r1 = s.draw();
            x = shell_position(d, ix, r1);
            vx = (VXA(ix) * (XA(ix + 1) - x) + VXA(ix + 1) * (x - XA(ix))) / DXX(ix);
            mu = (mu + vx * CINV) / (1.0 + vx * mu * CINV);
            wl = wl * (1.0 - vx * mu * CINV);
            const double help = 1.0 / (1.0 - vx * mu * CINV);
            evelo = evelo + e * (1.0 - help);
            e = e * help;
            e0 = e0 * help;
        }
    } else if (r1 >= 0.0 + pdop && r1 < 0.0 + pdop + pl) {
        // left leakage
        p.idist = -3;
        if (ix == 1) goto ret_early;
        const int l = ix - 1;
        if (glump == 0) iiig = ig;
        else {
            r1 = s.draw();
            const int n = cc.nlump;
            int j = guided_first_above(d.cumL + row, n, d.guide + ((size_t)0 * nx + ic - 1) * GUIDE_K, d.guide_m[0 * nx + ic - 1], r1);
            iiig = d.gorder[row + (j < n ? j : n - 1)];
        }
        if ((SIGT(l) + CAPT(iiig, l)) * DXX(ix - 1) < d.tauddmc) {
            p.itype = 1;
            r1 = s.draw();
            wl = 1.0 / (r1 * WLI(iiig + 1) + (1.0 - r1) * WLI(iiig));
            x = XA(ix);
            vx = VXA(ix);
            r1 = s.draw();
            r2 = s.draw();
            mu = -(r1 > r2 ? r1 : r2);
            mu = (mu + vx * CINV) / (1.0 + vx * mu * CINV);
            const double help = 1.0 / (1.0 - vx * mu * CINV);
            evelo = evelo + e * (1.0 - help);
            e = e * help;
            e0 = e0 * help;
            wl = wl * (1.0 - vx * mu * CINV);
        }
        ix = ix - 1;
        ig = iiig;
    } else if (r1 >= 0.0 + pdop + pl && r1 < 0.0 + pdop + (pl + pr)) {
        // right leakage
        p.idist = -4;
        if (ix == nx) {
            p.stat = 1;
            r1 = s.draw();
            r2 = s.draw();
            mu = r1 > r2 ? r1 : r2;
            if (glump != 0) {
                r1 = s.draw();
                const int n = cc.nlump;
                int j = guided_first_above(d.cumR + row, n, d.guide + ((size_t)1 * nx + ic - 1) * GUIDE_K, d.guide_m[1 * nx + ic - 1], r1);
                ig = d.gorder[row + (j < n ? j : n - 1)];
            }
            r1 = s.draw();
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
            goto ret_early;
        }
        const int l = ix + 1;
        if (glump == 0) iiig = ig;
        else {
            r1 = s.draw();
            const int n = cc.nlump;
            int j = guided_first_above(d.cumR + row, n, d.guide + ((size_t)1 * nx + ic - 1) * GUIDE_K, d.guide_m[1 * nx + ic - 1], r1);
            iiig = d.gorder[row + (j < n ? j : n - 1)];
        }
        if ((SIGT(l) + CAPT(iiig, l)) * DXX(ix + 1) < d.tauddmc) {
            p.itype = 1;
            r1 = s.draw();
            wl = 1.0 / (r1 * WLI(iiig + 1) + (1.0 - r1) * WLI(iiig));
            x = XA(ix + 1);
            vx = VXA(ix + 1);
            r1 = s.draw();
            r2 = s.draw();
            mu = r1 > r2 ? r1 : r2;
            mu = (mu + vx * CINV) / (1.0 + vx * mu * CINV);
            const double help = 1.0 / (1.0 - vx * mu * CINV);
            evelo = evelo + e * (1.0 - help);
            e = e * help;
            e0 = e0 * help;
            wl = wl * (1.0 - vx * mu * CINV);
        }
        ix = ix + 1;
        ig = iiig;
    } else {
        // effective scattering
        p.idist = -2;
        if (glump == ng) { ret = 102; goto ret_early; }
        if (glump == 0) {
            r1 = s.draw();
            iiig = emitgroup(d, r1, ic);
        } else {
            r1 = s.draw();
            const int n = cc.nlump;
            int j = n + guided_first_above(d.cumS + row + n, ng - n, d.guide + ((size_t)3 * nx + ic - 1) * GUIDE_K, d.guide_m[3 * nx + ic - 1], r1);
            iiig = d.gorder[row + (j < ng ? j : ng - 1)];
        }
        ig = iiig;
        if ((SIGT(ic) + CAPT(ig, ic)) * dist < d.tauddmc) {
            p.itype = 1;
            r1 = s.draw();
            wl = 1.0 / ((1.0 - r1) * WLI(ig) + r1 * WLI(ig + 1));
            r1 = s.draw();
            mu = 1.0 - 2.0 * r1;
            r1 = s.draw();
            x = shell_position(d, ix, r1);
            vx = (VXA(ix) * (XA(ix + 1) - x) + VXA(ix + 1) * (x - XA(ix))) / DXX(ix);
            mu = (mu + vx * CINV) / (1.0 + vx * mu * CINV);
            const double help = 1.0 / (1.0 - vx * mu * CINV);
            evelo = evelo + e * (1.0 - help);
            e = e * help;
            e0 = e0 * help;
            wl = wl * (1.0 - vx * mu * CINV);
        }
    }
    if (x < XA(1)) {
        p.stat = 2;
        goto ret_early;
    }
    {
        const double f = exp(-DIVV(ix) * ddmct / 3);
        evelo = evelo + e * (1.0 - f);
        e = e * f;
        e0 = e0 * f;
    }
ret_early:
    p.x = x; p.mu = mu; p.e = e; p.e0 = e0; p.wl = wl; p.vx = vx;
    p.ix = ix; p.ig = ig;
    return ret;
}

// particle_advance.f90:140-294 as a persistent kernel; see the file header.
#ifndef SLGPU_TRANSPORT_WAVES
#define SLGPU_TRANSPORT_WAVES 2
#endif
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SLGPU_TRANSPORT_WAVES, 8)))
void k_transport(Dev d, uint32_t iteration)
{
    extern __shared__ double s_tally[];
    const int ng = d.ng, nx = d.nx;
    for (int k = threadIdx.x; k < nx; k += blockDim.x) s_tally[k] = 0.0;
    __syncthreads();

    double evelo = 0.0, eout = 0.0, edead = 0.0, estart = 0.0, esrc = 0.0;
    unsigned long long nact = 0, nimc = 0, nddmc = 0, nswap = 0, nflux = 0, ndead = 0,
                       nerr = 0, nwarn = 0, nmax = 0;
    int run_ic = -1, run_ig = -1;
    double run_e = 0.0, run_j = 0.0;

    Ptcl p;
    Philox s;
    int64_t ip = -1, istep = 0;
    bool have = false;
    for (;;) {
        if (!have) {
            ip = (int64_t)atomicAdd(d.queue, 1ull);
            if (ip >= d.npart) break;
            // particle_advance.f90:145-202
            p.x = d.px[ip]; p.mu = d.pmu[ip]; p.e = d.pe[ip]; p.e0 = d.pe0[ip]; p.wl = d.pwl[ip];
            esrc += p.e;
            p.ix = binsrch(p.x, d.xarr, nx + 1, false);
            p.ig = binsrch(p.wl, d.wl, ng + 1, false);
            bool bad = p.ix < 1 || p.ig < 1;     // NaN position or wavelength
            if (bad) { p.ix = 1; p.ig = 1; }
            const double help = DXX(p.ix);
            const double tau = (SIGT(p.ix) + CAPT(p.ig, p.ix)) * help;
            p.itype = (d.puretran || tau < d.tauddmc) ? 1 : 2;
            const double h1 = XA(p.ix), h2 = XA(p.ix + 1);
            p.vx = (VXA(p.ix) * (h2 - p.x) + VXA(p.ix + 1) * (p.x - h1)) / (h2 - h1);
            if (p.itype == 1) {
                const double labfact = 1.0 - p.vx * p.mu / PC_C;
                p.wl = p.wl * labfact;
                p.e = p.e / labfact;
                p.e0 = p.e0 / labfact;
            }
            estart += p.e;
            p.stat = 0;
            p.idist = 0;
            s.init(d.seed, iteration, (uint32_t)ip, 1u);
            istep = 0;
            nact++;
            have = true;
            if (bad) {
                nerr++;
                ndead++;
                edead += p.e;
                d.pstat[ip] = 2;
                d.perr[ip] = 1;
                d.pnstep[ip] = 0;
                have = false;
                continue;
            }
        }
        // one event
        istep++;
        const int icold = p.ix, igold = p.ig;
        double erd, jrd;
        int ierr;
        if (p.itype == 1 || d.puretran) {
            nimc++;
            ierr = transport11(d, p, s, erd, jrd, evelo);
            if (p.itype != 1) nswap++;
        } else {
            nddmc++;
            ierr = diffusion11(d, p, s, erd, jrd, evelo);
            if (p.itype == 1) nswap++;
        }
        if (icold != run_ic || igold != run_ig) {
            if (run_ic > 0) {
                atomicAdd(&s_tally[run_ic - 1], run_e);
                atomicAdd(&d.jrad[(size_t)(run_ic - 1) * ng + run_ig - 1], run_j);
            }
            run_ic = icold; run_ig = igold; run_e = 0.0; run_j = 0.0;
        }
        run_e += erd;
        run_j += jrd;
        if (p.itype == 1 && p.stat == 0) {
            if (p.x > XA(p.ix + 1) || p.x < XA(p.ix) || p.x != p.x) {
                if (ierr == 0) ierr = -99;
            }
        }
        if (ierr < 0) nwarn++;
        if (ierr == 0 && p.stat == 0 && istep >= d.max_steps) ierr = 200;
        if (ierr > 0) {
            nerr++;
            p.stat = 2;
        }
        if (p.stat != 0) {
            // particle_advance.f90:271-292
            if ((unsigned long long)istep > nmax) nmax = istep;
            if (p.stat == 2) {
                ndead++;
                edead += p.e;
            } else {
                nflux++;
                eout += p.e;
            }
            d.pstat[ip] = (uint8_t)p.stat;
            d.px[ip] = p.x; d.pmu[ip] = p.mu; d.pwl[ip] = p.wl; d.pe[ip] = p.e; d.pe0[ip] = p.e0;
            d.pig[ip] = p.ig;
            d.pnstep[ip] = istep;
            d.perr[ip] = ierr > 0 ? ierr : 0;
            have = false;
        }
    }
    if (run_ic > 0) {
        atomicAdd(&s_tally[run_ic - 1], run_e);
        atomicAdd(&d.jrad[(size_t)(run_ic - 1) * ng + run_ig - 1], run_j);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nx; k += blockDim.x)
        if (s_tally[k] != 0.0) atomicAdd(&d.tally[k], s_tally[k]);

    // block reductions of totals and counters
    const double vals[5] = {eout, evelo, estart, edead, esrc};
    const int slot[5] = {0, 1, 3, 4, 5};
    for (int q = 0; q < 5; ++q) {
        double v = wave_sum(vals[q]);
        if ((threadIdx.x & 63) == 0) atomicAdd(&d.tot[slot[q]], v);
    }
    const unsigned long long cv[8] = {nact, nimc, nddmc, nswap, nflux, ndead, nerr, nwarn};
    const int cslot[8] = {0, 1, 2, 4, 5, 6, 7, 8};
    for (int q = 0; q < 8; ++q) {
        unsigned long long v = cv[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(&d.ct[cslot[q]], v);
    }
    {
        unsigned long long v = nmax;
        for (int off = 32; off > 0; off >>= 1) {
            unsigned long long o = __shfl_xor(v, off, 64);
            v = o > v ? o : v;
        }
        if ((threadIdx.x & 63) == 0) atomicMax(&d.ct[3], v);
    }
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
    int64_t need = (c->d.npart + 255) / 256;
    if (need < blocks) blocks = need > 0 ? need : 1;
    hipLaunchKernelGGL(k_transport, dim3((unsigned)blocks), dim3(256), c->nx * sizeof(double),
                       c->stream, c->d, iteration);
}

void launch_fluxtally(slgpu_ctx *c)
{
    int64_t n = c->d.npart;
    if (n == 0) return;
    hipLaunchKernelGGL(k_fluxtally, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       c->d);
}

}  // namespace slgpu
