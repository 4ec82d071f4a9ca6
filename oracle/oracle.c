/* ORACLE — CPU FP64 restatement of SuperLite's per-iteration hot path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Each function cites the reference
 * file:line it restates; expression order follows the reference's Fortran so that
 * FP64 results agree with it to a few ulp.
 */
#include "oracle.h"
#include "philox.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* physconstmod.f:11-51 */
#define PC_ACOEF 7.5657e-15
#define PC_PI 3.1415926535897932385
#define PC_PI2 (2.0 * PC_PI)
#define PC_PI4 (4.0 * PC_PI)
#define PC_C 2.997924562e10
#define PC_KB 1.3806505e-16
#define PC_ME 9.1093826e-28
#define PC_H 6.6260693e-27
#define PC_E 4.80325e-10
#define PC_RYDBERG 1.0973731534e5
#define PC_DEXT 0.7104
#define PC_EV 1.60217653e-12
#define PC_ANG 1e-8
#define PC_MBARN 1e-18
#define PC_AMU 1.66053886e-24

static const double CINV = 1.0 / PC_C;

#define FF_NU 81
#define FF_NGG 41

/* ------------------------------------------------------------------------ */
/* MISC/binsrch.f90:3-77 — interval index of x in arr(1:n), 1-based result.  */
int or_binsrch(double x, const double *arr, int n, int widerange)
{
    if (n == 2) {
        int r;
        if (x >= arr[0] && x <= arr[1]) r = 1;
        else if (x < arr[0]) r = 0;
        else if (x > arr[1]) r = 2;
        else return -1;
        if (!widerange) r = 1;
        return r;
    }
    int imin = 1, imax = n, imid = (n + 1) / 2;
    while (imax - imin > 1) {
        if (x >= arr[imin - 1] && x < arr[imid - 1]) {
            imax = imid;
            imid = (imax + imin) / 2;
        } else if (x >= arr[imid - 1] && x <= arr[imax - 1]) {
            imin = imid;
            imid = (imax + imin) / 2;
        } else {
            if (x < arr[0]) { imin = 0; imid = 0; break; }
            else if (x > arr[n - 1]) { imin = n; imid = n; break; }
            else return -1;
        }
    }
    if (imid != imin) imid = -1;
    if (!widerange) {
        if (imid > n - 1) imid = n - 1;
        if (imid < 1) imid = 1;
    }
    return imid;
}

/* ------------------------------------------------------------------------ */
/* groupmod.f:146-272 specintv, mode 0 (midpoint in x = hc/(wl k T)); offset 1-based */
static const double HCK = PC_H * PC_C / PC_KB;
static const double FTPI4 = 15.0 / (PC_PI * PC_PI * PC_PI * PC_PI);

static double fplanck(double x)
{
    if (x >= 300.0) return 300.0 * 300.0 * 300.0 / (exp(300.0) - 1.0);
    return x * x * x / (exp(x) - 1.0);
}

void or_specintv(double tempinv, const double *wlinv, int n, int offset, double *ss)
{
    double a = HCK * tempinv;
    double x0 = a * wlinv[offset - 1];
    for (int k = 0; k < n; ++k) {
        double x1 = a * wlinv[offset + k];
        ss[k] = FTPI4 * fabs(x1 - x0) * fplanck(0.5 * (x1 + x0));
        x0 = x1;
    }
}

/* groupmod.f:275-404 spectintv, mode 0, with the NaN -> 0 guard of :390-392 */
static double fdplanck(double x)
{
    if (x >= 300.0) {
        double e = exp(300.0);
        return (300.0 * 300.0 * 300.0 * 300.0 * e) / ((e - 1.0) * (e - 1.0));
    }
    double e = exp(x);
    return (x * x * x * x * e) / ((e - 1.0) * (e - 1.0));
}

void or_spectintv(double tempinv, const double *wlinv, int n, double *ss)
{
    const double ftpi4t = 15.0 / (4.0 * PC_PI * PC_PI * PC_PI * PC_PI);
    double a = HCK * tempinv;
    double x0 = a * wlinv[0];
    for (int k = 0; k < n; ++k) {
        double x1 = a * wlinv[k + 1];
        double v = ftpi4t * fabs(x1 - x0) * fdplanck(0.5 * (x1 + x0));
        ss[k] = (v != v) ? 0.0 : v;
        x0 = x1;
    }
}

/* groupmod.f:99-120 */
double or_specint0(double tempinv, const double *wlinv, int ig)
{
    double x = HCK * tempinv;
    double dx = x * fabs(wlinv[ig] - wlinv[ig - 1]);
    x = x * 0.5 * (wlinv[ig] + wlinv[ig - 1]);
    if (x >= 300.0) x = 300.0;
    return FTPI4 * dx * (x * x * x) / (exp(x) - 1.0);
}

/* groupmod.f:123-143 */
double or_dopspeccalc(double tempinv, const double *wlinv, int ig)
{
    double x = HCK * tempinv;
    x = x * wlinv[ig];
    if (x >= 300.0) x = 300.0;
    return FTPI4 * (x * x * x * x) / (exp(x) - 1.0);
}

/* physconstmod.f:79-93 */
double or_planck(double wl, double temp)
{
    const double c1 = 2.0 * PC_H * PC_C * PC_C;
    const double c2 = PC_H * PC_C / PC_KB;
    double b = 1.0 / (wl * temp);
    if (c2 * b >= 300.0) b = 300.0 / c2;
    double tb = temp * b;
    return c1 * (tb * tb * tb * tb * tb) / (exp(c2 * b) - 1.0);
}

/* ------------------------------------------------------------------------ */
/* bfxsmod.f:65-153 (Verner 1995/96 fits); ph1(6,7,30,30), ph2(7,30,30) Fortran order */
#define PH1(p, s, z, e) ((double)ph1[(p - 1) + 6 * ((s - 1) + 7 * ((z - 1) + 30 * (e - 1)))])
#define PH2(p, z, e) ((double)ph2[(p - 1) + 7 * ((z - 1) + 30 * (e - 1))])
/* bf_ph1(6,..)**2, bf_ph2(5,..)**2, bf_ph2(7,..)**2 are real*4 expressions in bfxsmod.f:141,148,149:
 * the square is rounded to single precision before it joins the double-precision sum */
#define SQ32(v) ((double)((float)(v) * (float)(v)))

double or_bfxs(int nz, int ne, double e, const float *ph1, const float *ph2)
{
    static const int ll[7] = {0, 0, 1, 0, 1, 2, 0};
    static const int ninn[30] = {0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 3, 3, 3, 3, 3, 3, 3, 3,
                                 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5};
    static const int ntot[30] = {1, 1, 2, 2, 3, 3, 3, 3, 3, 3, 4, 4, 5, 5, 5, 5, 5, 5,
                                 6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 7, 7};
    double xs = 0.0;
    if (nz < 1 || nz > 30) return xs;
    if (ne < 1 || ne > nz) return xs;
    int nout = ntot[ne - 1];
    if (nz > 18) {
        if (nz == ne) nout = 7;
        if (nz == ne + 1) {
            if (nz == 20 || nz == 21 || nz == 22 || nz == 25 || nz == 26) nout = 7;
        }
    }
    if (e < PH1(1, nout, nz, ne)) return xs;
    int nin = ninn[ne - 1];
    double einn;
    if (nz == 15 || nz == 17 || nz == 19 || (nz > 20 && nz != 26)) {
        einn = 0.0;
    } else {
        if (ne < 3) einn = (double)1.0e30f;
        else einn = PH1(1, nin, nz, ne);
    }
    xs = 0.0;
    for (int is = 1; is <= nout; ++is) {
        if (e < PH1(1, is, nz, ne)) continue;
        if (is < nout && is > nin && e < einn) continue;
        if (is <= nin || e >= einn) {
            double p1 = -PH1(5, is, nz, ne);
            double y = e / PH1(2, is, nz, ne);
            double q = -0.5 * p1 - ll[is - 1] - 5.5;
            double a = PH1(3, is, nz, ne) * ((y - 1.0) * (y - 1.0) + SQ32(PH1(6, is, nz, ne)));
            double b = sqrt(y / PH1(4, is, nz, ne)) + 1.0;
            xs = xs + a * pow(y, q) * pow(b, p1);
        } else {
            double p1 = -PH2(4, nz, ne);
            double q = -0.5 * p1 - 5.5;
            double x = e / PH2(1, nz, ne) - PH2(6, nz, ne);
            double z = sqrt(x * x + SQ32(PH2(7, nz, ne)));
            double a = PH2(2, nz, ne) * ((x - 1.0) * (x - 1.0) + SQ32(PH2(5, nz, ne)));
            double b = 1.0 + sqrt(z / PH2(3, nz, ne));
            xs = xs + a * pow(z, q) * pow(b, p1);
        }
    }
    return xs;
}

/* ------------------------------------------------------------------------ */
/* GAS/physical_opacity.f:3-357 (grp_ngs = 1) followed by opacity_planckmean.f:18-24
 * and opacity_rossmean.f:18-21.  cap64 is the FP64 `cap` before the sngl() rounding of
 * physical_opacity.f:267, returned group-fastest like gas_cap. */
int or_opacity(const or_opac_in *in, double *cap64, float *cap32, double *sig,
               double *capgrey, double *capross, float *emiss)
{
    const int nc = in->nc, ng = in->ng;
    const double *wlsub = in->wl;
    const double cthomson = 8.0 * PC_PI * (PC_E * PC_E * PC_E * PC_E) /
                            (3.0 * (PC_ME * PC_ME) * (PC_C * PC_C * PC_C * PC_C));
    const double c1 = 4.0 * (PC_E * PC_E * PC_E * PC_E * PC_E * PC_E) /
                      (3.0 * PC_H * PC_ME * (PC_C * PC_C * PC_C * PC_C)) *
                      sqrt(PC_PI2 / (3 * PC_ME * PC_H * PC_C));
    double *hckt = malloc(sizeof(double) * nc);
    double *expfac = malloc(sizeof(double) * nc);
    double *cap = calloc((size_t)nc * ng, sizeof(double));   /* (nc, ng) cell-fastest as :47 */
    /* ion lookup (iz, ii) -> k */
    int lut[31][32];
    for (int a = 0; a < 31; ++a)
        for (int b = 0; b < 32; ++b) lut[a][b] = -1;
    for (int k = 0; k < in->nion; ++k) lut[in->ion_iz[k]][in->ion_ii[k]] = k;

    for (int i = 0; i < nc; ++i) hckt[i] = PC_H * PC_C / (PC_KB * in->temp[i]);   /* :81 */
    for (int i = 0; i < nc; ++i) {                                               /* :84-89 */
        if (in->flags & 8) sig[i] = 0.0;
        else sig[i] = cthomson * in->nelec[i] * in->natom[i] / in->vol[i];
        if (in->mass[i] <= 0.0) sig[i] = 0.0;
    }

    /* bound-bound, :94-158 */
    if (!(in->flags & 1)) {
        int iigs = 0, dirty = 1;
        double phi = 0.0;
        for (int il = 0; il < in->nline; ++il) {
            double wl0 = (double)in->wl0[il] * PC_ANG;
            int igs;
            for (igs = iigs; igs <= ng; ++igs) {        /* pointer walk :118-121 */
                if (wlsub[igs] > wl0) break;            /* wlsub(igs+1) */
                dirty = 1;
            }
            iigs = igs;
            if (igs < 1) continue;
            if (igs > ng) continue;
            if (dirty) {
                dirty = 0;
                double wl = 0.5 * (wlsub[igs - 1] + wlsub[igs]);
                double wlinv = 1.0 / wl;
                double dwl = wlsub[igs] - wlsub[igs - 1];
                for (int i = 0; i < nc; ++i) expfac[i] = 1.0 - exp(-hckt[i] * wlinv);
                phi = wl * wl / (dwl * PC_C);
            }
            int k = lut[in->liz[il]][in->lii[il]];
            if (k < 0) continue;
            double gxs = (double)in->gxs[il], chilw = (double)in->chilw[il];
            for (int i = 0; i < nc; ++i) {
                if (in->mass[i] <= 0.0) continue;
                double ocg = in->oc[k + (size_t)in->nion * i] * in->ginv[k + (size_t)in->nion * i];
                if (ocg <= 0.0) continue;
                double caphelp = phi * gxs * ocg * exp(-(chilw * hckt[i])) * expfac[i];
                cap[i + (size_t)nc * (igs - 1)] += caphelp;
            }
        }
    }

    /* bound-free, :163-195 (left group edge, photon energy in eV) */
    if (!(in->flags & 2)) {
        for (int igs = 1; igs <= ng; ++igs) {
            double wl = wlsub[igs - 1];
            double en = PC_H * PC_C / (PC_EV * wl);
            for (int k = 0; k < in->nion; ++k) {
                int iz = in->ion_iz[k], ii = in->ion_ii[k];
                int ie = iz - ii + 1;
                double xs = or_bfxs(iz, ie, en, in->bf_ph1, in->bf_ph2);
                if (xs == 0.0) continue;
                for (int i = 0; i < nc; ++i) {
                    if (in->mass[i] <= 0.0) continue;
                    cap[i + (size_t)nc * (igs - 1)] += xs * PC_MBARN * in->oc[k + (size_t)in->nion * i];
                }
            }
        }
    }

    /* free-free, :200-247 (Sutherland gaunt factor, bilinear in log u and log gamma^2) */
    if (!(in->flags & 4)) {
        for (int igs = 1; igs <= ng; ++igs) {
            double wl = wlsub[igs - 1];
            double wlinv = 1.0 / wl;
            for (int i = 0; i < nc; ++i) {
                if (in->mass[i] <= 0.0) continue;
                double nv = in->natom[i] / in->vol[i];
                double hlp = nv * nv * in->nelec[i];
                double u = hckt[i] * wlinv;
                double ru = 10.0 * (log10(u) + 4.0) + 1.0;
                int iu = (int)floor(ru);
                if (iu < 1) iu = 1;
                if (iu > FF_NU - 1) iu = FF_NU - 1;
                double du = ru - iu;
                double help = c1 * sqrt(hckt[i]) * (1.0 - exp(-u)) * (wl * wl * wl) * hlp;
                for (int iz = 1; iz <= 30; ++iz) {
                    double gg = (double)(iz * iz) * PC_RYDBERG * hckt[i];
                    double rgg = 5.0 * (log10(gg) + 4.0) + 1.0;
                    int igg = (int)floor(rgg);
                    if (igg < 1) igg = 1;
                    if (igg > FF_NGG - 1) igg = FF_NGG - 1;
                    double dgg = rgg - igg;
                    const double *T = in->ff_gff;
#define GFF(a, b) T[((b) - 1) * FF_NU + ((a) - 1)]
                    double gff = (1.0 - du) * (1.0 - dgg) * GFF(iu, igg) +
                                 du * (1.0 - dgg) * GFF(iu + 1, igg) +
                                 (1.0 - du) * dgg * GFF(iu, igg + 1) +
                                 du * dgg * GFF(iu + 1, igg + 1);
#undef GFF
                    if (rgg > FF_NGG) gff = gff > 1.0 ? gff : 1.0;
                    cap[i + (size_t)nc * (igs - 1)] += help * gff * (double)(iz * iz) *
                                                       in->natomfr[(iz - 1) + 30 * (size_t)i];
                }
            }
        }
    }

    /* store (:267) and emissivity (:269-279) */
    for (int i = 0; i < nc; ++i) {
        for (int ig = 1; ig <= ng; ++ig) {
            double c = cap[i + (size_t)nc * (ig - 1)];
            if (cap64) cap64[(ig - 1) + (size_t)ng * i] = c;
            cap32[(ig - 1) + (size_t)ng * i] = (float)c;
            if (emiss) {
                double em = 0.0;
                if (in->mass[i] > 0.0) {
                    double wl = 0.5 * (in->wl[ig] + in->wl[ig - 1]);
                    double bnu = (wl * wl / PC_C) * or_planck(wl, in->temp[i]);
                    em = c * bnu;
                }
                emiss[(ig - 1) + (size_t)ng * i] = (float)em;
            }
        }
    }

    /* Planck and Rosseland means of the FP32 group opacity */
    double *wlinv = malloc(sizeof(double) * (ng + 1));
    double *ss = malloc(sizeof(double) * ng);
    for (int g = 0; g <= ng; ++g) wlinv[g] = 1.0 / in->wl[g];
    for (int i = 0; i < nc; ++i) {
        or_specintv(1.0 / in->temp[i], wlinv, ng, 1, ss);
        double s = 0.0;
        for (int g = 0; g < ng; ++g) s += (double)cap32[g + (size_t)ng * i] * ss[g];
        capgrey[i] = s;
        or_spectintv(1.0 / in->temp[i], wlinv, ng, ss);
        s = 0.0;
        for (int g = 0; g < ng; ++g) s += (1.0 / (sig[i] + (double)cap32[g + (size_t)ng * i])) * ss[g];
        capross[i] = 1.0 / s;
    }
    free(wlinv); free(ss);
    free(hckt); free(expfac); free(cap);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* LAPACK dgetf2 (unblocked LU with partial pivoting, row interchanges) and dgetrs (no
 * transpose) on a column-major n x n matrix; the updates are written as fma so that the
 * device restatement can reproduce them operation for operation. */
#define A_(i, j) a[(size_t)(j) * n + (i)]
int or_lu_solve(int n, double *a, int *ipiv, double *b)
{
    int info = 0;
    const double sfmin = DBL_MIN;
    for (int j = 0; j < n; ++j) {
        int jp = j;
        double amax = fabs(A_(j, j));
        for (int i = j + 1; i < n; ++i)                 /* idamax: first largest |a| */
            if (fabs(A_(i, j)) > amax) { amax = fabs(A_(i, j)); jp = i; }
        ipiv[j] = jp;
        if (A_(jp, j) != 0.0) {
            if (jp != j)
                for (int k = 0; k < n; ++k) { double t = A_(j, k); A_(j, k) = A_(jp, k); A_(jp, k) = t; }
            if (fabs(A_(j, j)) >= sfmin) {
                const double r = 1.0 / A_(j, j);
                for (int i = j + 1; i < n; ++i) A_(i, j) = A_(i, j) * r;
            } else {
                for (int i = j + 1; i < n; ++i) A_(i, j) = A_(i, j) / A_(j, j);
            }
        } else if (info == 0) {
            info = j + 1;
        }
        for (int k = j + 1; k < n; ++k) {               /* dger: a -= l u */
            const double u = A_(j, k);
            if (u != 0.0)
                for (int i = j + 1; i < n; ++i) A_(i, k) = fma(-A_(i, j), u, A_(i, k));
        }
    }
    if (info) return info;
    for (int k = 0; k < n; ++k)                          /* dlaswp */
        if (ipiv[k] != k) { double t = b[k]; b[k] = b[ipiv[k]]; b[ipiv[k]] = t; }
    for (int k = 0; k < n; ++k)                          /* dtrsm lower, unit */
        if (b[k] != 0.0)
            for (int i = k + 1; i < n; ++i) b[i] = fma(-b[k], A_(i, k), b[i]);
    for (int k = n - 1; k >= 0; --k)                     /* dtrsm upper, non-unit */
        if (b[k] != 0.0) {
            b[k] = b[k] / A_(k, k);
            for (int i = 0; i < k; ++i) b[i] = fma(-b[k], A_(i, k), b[i]);
        }
    return 0;
}
#undef A_

/* physical_opacity_nlte.f:148-263: the H I level populations of cell i (normalised to 1) */
static int or_nlte_pops(const or_opac_in *in, const or_nlte_in *nl, int i, double *rc,
                        double *a, int *ipiv, double *gpi, double *rhs)
{
    const int n = nl->nlev, ng = in->ng;
    const double *wlsub = in->wl;
    const double kb_ev = PC_KB / PC_EV;
    const double tempinv = 1.0 / in->temp[i];
#define RC(l, ll) rc[(size_t)((ll) - 1) * n + ((l) - 1)]
    for (size_t k = 0; k < (size_t)n * n; ++k) rc[k] = 0.0;
    for (int l = 0; l < n; ++l) rhs[l] = 0.0;
    rhs[0] = 1.0;
    /* radiative terms, :160-184 (J at the group centre, Planck at the gas temperature) */
    for (int il = 0; il < nl->nlin; ++il) {
        const double wl0 = nl->wl0[il] * PC_ANG;
        int igs;
        for (igs = 0; igs <= ng; ++igs)
            if (wlsub[igs] > wl0) break;                 /* wlsub(igs+1) > wl0 */
        if (igs < 1 || igs > ng) continue;
        const double wlm = 0.5 * (wlsub[igs] + wlsub[igs - 1]);
        const double jnu = (wlm * wlm / PC_C) * or_planck(wlm, in->temp[i]);
        RC(nl->llw[il], nl->lup[il]) = nl->blu[il] * jnu;
        RC(nl->lup[il], nl->llw[il]) = nl->aul[il] + nl->bul[il] * jnu;
    }
    /* electron-impact excitation, :189-222 */
    for (int il = 0; il < nl->ncoll; ++il) {
        const double x = nl->dele[il] * tempinv / kb_ev;
        const double lx = log(x);
        const double *c = nl->coll_c + 4 * (size_t)il;
        double f = c[0] + c[1] * lx + c[2] * (lx * lx) + c[3] * (lx * lx * lx);
        f = exp(f);
        const double gamma = 1.58e-5 * (f / x) * exp(-1.0 * x) * pow(tempinv, 1.5) / pow(kb_ev, 1.5);
        const int lw = nl->llw_coll[il], lu = nl->lup_coll[il];
        const double gamma_d = nl->glev[lw - 1] * gamma * exp(x) / nl->glev[lu - 1];
        const double clu = gamma * in->nelec[i] * in->natom[i] / in->vol[i];
        const double cul = gamma_d * in->nelec[i] * in->natom[i] / in->vol[i];
        RC(lw, lu) = RC(lw, lu) + clu;
        RC(lu, lw) = RC(lu, lw) + cul;
    }
    /* photoionisation and radiative recombination, :226-253 */
    for (int ll = 0; ll < n; ++ll) {
        const int l = nl->lev_pi[ll];
        const double xpi = nl->dele_pi[ll] * tempinv / kb_ev;
        const double lp = log(xpi);
        const double *cp = nl->pi_c + 4 * (size_t)ll, *cr = nl->rr_c + 4 * (size_t)ll;
        double fpi = cp[0] + cp[1] * lp + cp[2] * (lp * lp) + cp[3] * (lp * lp * lp);
        fpi = exp(fpi);
        gpi[l - 1] = (fpi / xpi) * exp(-1.0 * xpi) * pow(tempinv, 1.5) / pow(kb_ev, 1.5);
        const double xrr = kb_ev * nl->radtemp[i];
        const double lr = log(xrr);
        double frr = cr[0] + cr[1] * lr + cr[2] * (lr * lr) + cr[3] * (lr * lr * lr);
        frr = exp(frr);
        if (l > 1) rhs[l - 1] = -1.0 * frr * nl->ndens[2 * (size_t)i + 1] / nl->ndens[2 * (size_t)i];
    }
    /* rate matrix, :255-266 (row 1: the populations sum to 1) */
    for (int l = 1; l <= n; ++l)
        for (int ll = 1; ll <= n; ++ll) {
            double v;
            if (l == 1) v = 1.0;
            else if (l == ll) {
                double sum = 0.0;
                for (int k = 1; k <= n; ++k) sum += RC(l, k);
                v = -1.0 * sum;
                v = v - gpi[l - 1];
            } else v = RC(ll, l);
            a[(size_t)(ll - 1) * n + (l - 1)] = v;
        }
#undef RC
    return ipiv ? or_lu_solve(n, a, ipiv, rhs) : 0;
}

int or_nlte_rate_matrix(const or_opac_in *in, const or_nlte_in *nl, int i, double *a, double *b)
{
    const int n = nl->nlev;
    double *rc = malloc(sizeof(double) * n * n), *gpi = calloc(n, sizeof(double));
    or_nlte_pops(in, nl, i, rc, a, NULL, gpi, b);
    free(rc); free(gpi);
    return 0;
}

int or_opacity_nlte(const or_opac_in *in, const or_nlte_in *nl, double *cap64, float *cap32,
                    float *capemit32, double *sig, double *capgrey, double *capemitgrey,
                    double *capross, float *emiss, double *npop)
{
    const int nc = in->nc, ng = in->ng, n = nl->nlev;
    const double *wlsub = in->wl;
    const double cthomson = 8.0 * PC_PI * (PC_E * PC_E * PC_E * PC_E) /
                            (3.0 * (PC_ME * PC_ME) * (PC_C * PC_C * PC_C * PC_C));
    const double c1 = 4.0 * (PC_E * PC_E * PC_E * PC_E * PC_E * PC_E) /
                      (3.0 * PC_H * PC_ME * (PC_C * PC_C * PC_C * PC_C)) *
                      sqrt(PC_PI2 / (3 * PC_ME * PC_H * PC_C));
    const double fconst = PC_PI * (PC_E * PC_E) / (PC_ME * PC_C);
    double *hckt = malloc(sizeof(double) * nc);
    double *expfac = malloc(sizeof(double) * nc);
    double *cap = calloc((size_t)nc * ng, sizeof(double));
    double *capemit = calloc((size_t)nc * ng, sizeof(double));
    double *em = calloc((size_t)nc * ng, sizeof(double));
    double *pops = calloc((size_t)nc * n, sizeof(double));
    double *rc = malloc(sizeof(double) * n * n), *a = malloc(sizeof(double) * n * n);
    double *gpi = calloc(n, sizeof(double)), *rhs = malloc(sizeof(double) * n);
    int *ipiv = malloc(sizeof(int) * n);
    int rc_ = 0;
    int lut[31][32];
    for (int x = 0; x < 31; ++x)
        for (int y = 0; y < 32; ++y) lut[x][y] = -1;
    for (int k = 0; k < in->nion; ++k) lut[in->ion_iz[k]][in->ion_ii[k]] = k;
    for (int i = 0; i < nc; ++i) hckt[i] = PC_H * PC_C / (PC_KB * in->temp[i]);
    for (int i = 0; i < nc; ++i) {
        if (in->flags & 8) sig[i] = 0.0;
        else sig[i] = cthomson * in->nelec[i] * in->natom[i] / in->vol[i];
        if (in->mass[i] <= 0.0) sig[i] = 0.0;
    }
#define AT(arr, i, igs) arr[(i) + (size_t)nc * ((igs) - 1)]
    if (!(in->flags & 1)) {
        /* populations, :140-263 */
        for (int i = 0; i < nc && !rc_; ++i) {
            if (in->mass[i] <= 0.0) continue;
            if (or_nlte_pops(in, nl, i, rc, a, ipiv, gpi, rhs)) rc_ = 1;
            for (int l = 0; l < n; ++l) pops[(size_t)i * n + l] = rhs[l];
        }
        /* NLTE H lines, :304-345 */
        int dirty = 1;
        double phi = 0.0, wl = 0.0;
        for (int il = 0; il < nl->nlin; ++il) {
            const double wl0 = nl->wl0[il] * PC_ANG;
            int igs;
            for (igs = 0; igs <= ng; ++igs) {
                if (wlsub[igs] > wl0) break;
                dirty = 1;
            }
            if (igs < 1 || igs > ng) continue;
            if (dirty) {
                dirty = 0;
                wl = 0.5 * (wlsub[igs - 1] + wlsub[igs]);
                const double dwl = wlsub[igs] - wlsub[igs - 1];
                phi = wl * wl / (dwl * PC_C);
            }
            const int lw = nl->llw[il], lu = nl->lup[il];
            const double g_llw = nl->gl[il], g_lup = nl->gu[il];
            for (int i = 0; i < nc; ++i) {
                if (in->mass[i] <= 0.0) continue;
                const double npl = pops[(size_t)i * n + lw - 1] * nl->ndens[2 * (size_t)i];
                const double npu = pops[(size_t)i * n + lu - 1] * nl->ndens[2 * (size_t)i];
                const double caphelp = npl * fconst * nl->f[il] * phi * (1.0 - (npu * g_llw / (npl * g_lup)));
                AT(cap, i, igs) += caphelp;
                const double bnu = (wl * wl / PC_C) * or_planck(wl, in->temp[i]);
                const double emithelp = npu * nl->aul[il] * phi * PC_H * PC_C / (4.0 * PC_PI * wl0);
                AT(capemit, i, igs) += emithelp / bnu;
                AT(em, i, igs) += emithelp;
            }
        }
        /* LTE lines of the other elements, :349-401 */
        int iigs = 0;
        dirty = 1;
        phi = 0.0;
        for (int il = 0; il < in->nline; ++il) {
            const double wl0 = (double)in->wl0[il] * PC_ANG;
            int igs;
            for (igs = iigs; igs <= ng; ++igs) {
                if (wlsub[igs] > wl0) break;
                dirty = 1;
            }
            iigs = igs;
            if (igs < 1 || igs > ng) continue;
            if (dirty) {
                dirty = 0;
                wl = 0.5 * (wlsub[igs - 1] + wlsub[igs]);
                const double wlinv = 1.0 / wl;
                const double dwl = wlsub[igs] - wlsub[igs - 1];
                for (int i = 0; i < nc; ++i) expfac[i] = 1.0 - exp(-hckt[i] * wlinv);
                phi = wl * wl / (dwl * PC_C);
            }
            const int k = lut[in->liz[il]][in->lii[il]];
            if (k < 0) continue;
            const double gxs = (double)in->gxs[il], chilw = (double)in->chilw[il];
            for (int i = 0; i < nc; ++i) {
                if (in->mass[i] <= 0.0) continue;
                const double ocg = in->oc[k + (size_t)in->nion * i] * in->ginv[k + (size_t)in->nion * i];
                if (ocg <= 0.0) continue;
                const double caphelp = phi * gxs * ocg * exp(-chilw * hckt[i]) * expfac[i];
                AT(cap, i, igs) += caphelp;
                AT(capemit, i, igs) += caphelp;
                const double bnu = (wl * wl / PC_C) * or_planck(wl, in->temp[i]);
                AT(em, i, igs) += caphelp * bnu;
            }
        }
    }
    /* bound-free, :413-452 (left group edge) */
    if (!(in->flags & 2)) {
        for (int igs = 1; igs <= ng; ++igs) {
            const double wl = wlsub[igs - 1];
            const double en = PC_H * PC_C / (PC_EV * wl);
            for (int k = 0; k < in->nion; ++k) {
                const int iz = in->ion_iz[k], ii = in->ion_ii[k];
                const double xs = or_bfxs(iz, iz - ii + 1, en, in->bf_ph1, in->bf_ph2);
                if (xs == 0.0) continue;
                for (int i = 0; i < nc; ++i) {
                    if (in->mass[i] <= 0.0) continue;
                    const double o = in->oc[k + (size_t)in->nion * i];
                    AT(cap, i, igs) += xs * PC_MBARN * o;
                    AT(capemit, i, igs) += xs * PC_MBARN * o;
                    const double bnu = (wl * wl / PC_C) * or_planck(wl, in->temp[i]);
                    AT(em, i, igs) += xs * PC_MBARN * o * bnu;
                }
            }
        }
    }
    /* free-free, :458-511 (left group edge) */
    if (!(in->flags & 4)) {
        for (int igs = 1; igs <= ng; ++igs) {
            const double wl = wlsub[igs - 1];
            const double wlinv = 1.0 / wl;
            for (int i = 0; i < nc; ++i) {
                if (in->mass[i] <= 0.0) continue;
                const double nv = in->natom[i] / in->vol[i];
                const double hlp = nv * nv * in->nelec[i];
                const double u = hckt[i] * wlinv;
                const double ru = 10.0 * (log10(u) + 4.0) + 1.0;
                int iu = (int)floor(ru);
                if (iu < 1) iu = 1;
                if (iu > FF_NU - 1) iu = FF_NU - 1;
                const double du = ru - iu;
                const double help = c1 * sqrt(hckt[i]) * (1.0 - exp(-u)) * (wl * wl * wl) * hlp;
                for (int iz = 1; iz <= 30; ++iz) {
                    const double gg = (double)(iz * iz) * PC_RYDBERG * hckt[i];
                    const double rgg = 5.0 * (log10(gg) + 4.0) + 1.0;
                    int igg = (int)floor(rgg);
                    if (igg < 1) igg = 1;
                    if (igg > FF_NGG - 1) igg = FF_NGG - 1;
                    const double dgg = rgg - igg;
                    const double *T = in->ff_gff;
#define GFF(x, y) T[((y) - 1) * FF_NU + ((x) - 1)]
                    double gff = (1.0 - du) * (1.0 - dgg) * GFF(iu, igg) + du * (1.0 - dgg) * GFF(iu + 1, igg) +
                                 (1.0 - du) * dgg * GFF(iu, igg + 1) + du * dgg * GFF(iu + 1, igg + 1);
#undef GFF
                    if (rgg > FF_NGG) gff = gff > 1.0 ? gff : 1.0;
                    const double t = help * gff * (double)(iz * iz) * in->natomfr[(iz - 1) + 30 * (size_t)i];
                    AT(cap, i, igs) += t;
                    AT(capemit, i, igs) += t;
                    const double bnu = (wl * wl / PC_C) * or_planck(wl, in->temp[i]);
                    AT(em, i, igs) += t * bnu;
                }
            }
        }
    }
    /* save (:554-556) and the means */
    for (int i = 0; i < nc; ++i)
        for (int ig = 1; ig <= ng; ++ig) {
            const size_t o = (ig - 1) + (size_t)ng * i;
            if (cap64) cap64[o] = AT(cap, i, ig);
            cap32[o] = (float)AT(cap, i, ig);
            capemit32[o] = (float)AT(capemit, i, ig);
            if (emiss) emiss[o] = (float)AT(em, i, ig);
        }
#undef AT
    double *wlinv = malloc(sizeof(double) * (ng + 1));
    double *ss = malloc(sizeof(double) * ng);
    for (int g = 0; g <= ng; ++g) wlinv[g] = 1.0 / in->wl[g];
    for (int i = 0; i < nc; ++i) {
        or_specintv(1.0 / in->temp[i], wlinv, ng, 1, ss);
        double s1 = 0.0, s2 = 0.0;
        for (int g = 0; g < ng; ++g) {
            s1 += (double)cap32[g + (size_t)ng * i] * ss[g];
            s2 += (double)capemit32[g + (size_t)ng * i] * ss[g];
        }
        capgrey[i] = s1;
        capemitgrey[i] = s2;
        or_spectintv(1.0 / in->temp[i], wlinv, ng, ss);
        double s = 0.0;
        for (int g = 0; g < ng; ++g) s += (1.0 / (sig[i] + (double)cap32[g + (size_t)ng * i])) * ss[g];
        capross[i] = 1.0 / s;
    }
    if (npop) memcpy(npop, pops, sizeof(double) * (size_t)nc * n);
    free(wlinv); free(ss); free(hckt); free(expfac); free(cap); free(capemit); free(em);
    free(pops); free(rc); free(a); free(gpi); free(rhs); free(ipiv);
    return rc_;
}

/* ------------------------------------------------------------------------ */
#define CAP(ig, ic) ((double)m->cap[((ic) - 1) * (size_t)m->ng + ((ig) - 1)])
#define XA(i) (m->xarr[(i) - 1])
#define VXA(i) (m->vxarr[(i) - 1])
#define DX(i) (m->xarr[(i)] - m->xarr[(i) - 1])
#define DX3(i) (m->xarr[(i)] * m->xarr[(i)] * m->xarr[(i)] - m->xarr[(i) - 1] * m->xarr[(i) - 1] * m->xarr[(i) - 1])
#define SIG(ic) (m->sig[(ic) - 1])
#define OPL(k, ic) (m->opaclump[((ic) - 1) * 11 + ((k) - 1)])
#define WL(ig) (m->wl[(ig) - 1])
#define WLINV(ig) (m->wlinv[(ig) - 1])
#define DIVV(ix) (m->divv[(ix) - 1])
#define TINV(ic) (m->tempinv[(ic) - 1])
/* NLTE emission opacity (in_nlte) */
#define NLTE (m->capemit != NULL)
#define CAPE(ig, ic) ((double)m->capemit[((ic) - 1) * (size_t)m->ng + ((ig) - 1)])

/* GRID/leakage_opacity11.f90:30-171 (LTE: capemit == 0, so opaclump(9) == 0 and the capemit
 * loop never moves igemitmax) */
void or_leakage(const or_model *m, double *opaclump)
{
    const int ng = m->ng, nx = m->nx;
    double *specarr = malloc(sizeof(double) * ng);
    for (int i = 1; i <= nx; ++i) {
        int l = i;
        double *op = opaclump + (size_t)(l - 1) * 11;
        for (int k = 0; k < 11; ++k) op[k] = 0.0;
        int icnb1 = i - 1 < 1 ? 1 : i - 1;
        int icnb2 = i + 1 > nx ? nx : i + 1;
        double dist = DX(i);
        or_specintv(TINV(l), m->wlinv, ng, 1, specarr);
        double speclump = 0.0;
        for (int ig = 1; ig <= ng; ++ig)
            if (CAP(ig, l) * dist >= m->taulump && (SIG(l) + CAP(ig, l)) * dist >= m->tauddmc)
                speclump += specarr[ig - 1];
        speclump = speclump > 0.0 ? 1.0 / speclump : 0.0;
        op[6] = speclump;
        double caplump = 0.0, emitmax = 0.0;
        int igemitmax = 0;
        for (int ig = 1; ig <= ng; ++ig) {
            if (CAP(ig, l) * dist < m->taulump) continue;
            if ((SIG(l) + CAP(ig, l)) * dist < m->tauddmc) continue;
            double help = specarr[ig - 1] * CAP(ig, l);
            caplump += help;
            if (help > emitmax) { emitmax = help; igemitmax = ig; }
        }
        /* leakage_opacity11.f90:76-85 capemitlump (NLTE) */
        double capemitlump = 0.0;
        if (NLTE)
            for (int ig = 1; ig <= ng; ++ig) {
                if (CAP(ig, l) * dist < m->taulump) continue;
                if ((SIG(l) + CAP(ig, l)) * dist < m->tauddmc) continue;
                double help = specarr[ig - 1] * CAPE(ig, l);
                capemitlump += help;
                if (help > emitmax) { emitmax = help; igemitmax = ig; }
            }
        double doplump = 0.0;
        if (DIVV(l) >= 0) {
            for (int ig = 1; ig <= ng - 1; ++ig) {
                if (CAP(ig, l) * dist < m->taulump) continue;
                if (CAP(ig + 1, l) * dist >= m->taulump) continue;
                if ((SIG(l) + CAP(ig, l)) * dist < m->tauddmc) continue;
                doplump += or_dopspeccalc(TINV(l), m->wlinv, ig) * DIVV(l) / (3 * PC_C);
            }
        } else {
            for (int ig = ng; ig >= 2; --ig) {
                if (CAP(ig, l) * dist < m->taulump) continue;
                if (CAP(ig - 1, l) * dist >= m->taulump) continue;
                if ((SIG(l) + CAP(ig, l)) * dist < m->tauddmc) continue;
                doplump -= or_dopspeccalc(TINV(l), m->wlinv, ig - 1) * DIVV(l) / (3 * PC_C);
            }
        }
        op[7] = caplump;
        op[8] = capemitlump;
        op[9] = (double)igemitmax;
        op[10] = doplump;
        double dvol = 3.0 * m->vol[l - 1] / PC_PI4;
        for (int ig = 1; ig <= ng; ++ig) {
            if (CAP(ig, l) * dist < m->taulump) continue;
            if ((SIG(l) + CAP(ig, l)) * dist < m->tauddmc) continue;
            double specval = specarr[ig - 1];
            int lhelp = (i == 1) ? 1 : ((CAP(ig, icnb1) + SIG(icnb1)) * DX(i - 1) < m->tauddmc);
            if (lhelp) {
                double help = (CAP(ig, l) + SIG(l)) * DX(i);
                double ppl = 4.0 / (3.0 * help + 6.0 * PC_DEXT);
                op[0] += (specval * speclump) * 1.5 * ppl * (XA(i) * XA(i)) / dvol;
            } else {
                double help = (SIG(l) + CAP(ig, l)) * DX(i) + (SIG(icnb1) + CAP(ig, icnb1)) * DX(i - 1);
                op[0] += (specval * speclump) * 2.0 * (XA(i) * XA(i)) / (help * 3.0 * m->vol[l - 1] / PC_PI4);
            }
            lhelp = (i == nx) ? 1 : ((CAP(ig, icnb2) + SIG(icnb2)) * DX(i + 1) < m->tauddmc);
            if (lhelp) {
                double help = (CAP(ig, l) + SIG(l)) * DX(i);
                double ppr = 4.0 / (3.0 * help + 6.0 * PC_DEXT);
                op[1] += (specval * speclump) * 1.5 * ppr * (XA(i + 1) * XA(i + 1)) / dvol;
            } else {
                double help = (SIG(l) + CAP(ig, l)) * DX(i) + (SIG(icnb2) + CAP(ig, icnb2)) * DX(i + 1);
                op[1] += (specval * speclump) * 2.0 * (XA(i + 1) * XA(i + 1)) / (help * 3.0 * m->vol[l - 1] / PC_PI4);
            }
        }
    }
    free(specarr);
}

/* GRID/emission_probability.f90:39-62 (grp_ng > 1; grd_capemit and grd_capemitgrey in NLTE) */
void or_emission_probability(const or_model *m, double *emitprob)
{
    const int ng = m->ng;
    double *specarr = malloc(sizeof(double) * ng);
    for (int l = 1; l <= m->nx; ++l) {
        or_specintv(TINV(l), m->wlinv, ng, 1, specarr);
        int ig = 1;
        double help = 0.0;
        for (int iep = 1; iep <= m->nep - 1; ++iep) {
            int hi = iep * m->nepg < ng ? iep * m->nepg : ng;
            int nepg = hi - ig + 1;
            int igp1 = ig + nepg - 1;
            double s = 0.0;
            for (int g = ig; g <= igp1; ++g) s += specarr[g - 1] * (NLTE ? CAPE(g, l) : CAP(g, l));
            help = help + s;
            emitprob[(size_t)(l - 1) * m->nep + iep - 1] = help;
            ig = igp1 + 1;
        }
        emitprob[(size_t)(l - 1) * m->nep + m->nep - 1] = NLTE ? m->capemitgrey[l - 1] : m->capgrey[l - 1];
    }
    free(specarr);
}

/* GRID/emitgroup.f:3-54 */
int or_emitgroup(const or_model *m, double r, int ic)
{
    const double *ep = m->emitprob + (size_t)(ic - 1) * m->nep;
    double r1 = r * (NLTE ? m->capemitgrey[ic - 1] : m->capgrey[ic - 1]);
    int iep = or_binsrch(r1, ep, m->nep, 1);
    int ig = iep * m->nepg + 1;
    int igp1 = ig + m->nepg - 1 < m->ng ? ig + m->nepg - 1 : m->ng;
    int nepg = igp1 - ig + 1;
    double specval[nepg > 0 ? nepg : 1];
    if (nepg > 0) or_specintv(TINV(ic), m->wlinv, nepg, ig, specval);
    double emitprob = (iep == 0) ? 0.0 : ep[iep - 1];
    int l = 0;
    for (; ig <= igp1 - 1; ++ig) {
        l = l + 1;
        emitprob = emitprob + specval[l - 1] * (NLTE ? CAPE(ig, ic) : CAP(ig, ic));
        if (emitprob > r1) break;
    }
    return ig;
}

/* ------------------------------------------------------------------------ */
/* LTE EOS.  ionsmod.f:311-354 saha_nelec: per element the Saha recursion from stage istart
 * (restarted one stage higher while the last stage overflows; the stages below keep their
 * values), normalised; returns sum_z natomfr(z) * (mean charge of z). */
#define EOS_E(iz, ii) (ed->e[((iz) - 1) * 31 + (ii) - 1])
#define EOS_Q(iz, ii) (ed->q[((iz) - 1) * 31 + (ii) - 1])
#define EOS_G(iz, ii) (ed->g1[((iz) - 1) * 31 + (ii) - 1])
static int or_saha_nelec(const or_eos_data *ed, const double *natomfr, double kti,
                         double sahac2, double *n /* 30*31 */, double *nelec_new)
{
    double ne_sum = 0.0;
    for (int iz = 1; iz <= 30; ++iz) {
        const int nion = ed->ni[iz - 1];
        double *nz = n + (iz - 1) * 31;
        double help = 0.0;
        for (int istart = 1; istart <= nion - 1; ++istart) {
            nz[istart - 1] = 1.0;
            for (int ii = istart + 1; ii <= nion; ++ii) {
                help = sahac2 * exp(-kti * EOS_E(iz, ii - 1)) * nz[ii - 2] * EOS_Q(iz, ii) / EOS_Q(iz, ii - 1);
                nz[ii - 1] = help;
            }
            if (help < DBL_MAX) break;
        }
        double nsum = 0.0;
        for (int ii = 1; ii <= nion; ++ii) nsum += nz[ii - 1];
        if (nsum != nsum || nsum > DBL_MAX || nsum < DBL_MIN) return 1;   /* :338-349 stop */
        nsum = 1.0 / nsum;
        for (int ii = 1; ii <= nion; ++ii) nz[ii - 1] = nz[ii - 1] * nsum;
        double ne = 0.0;
        for (int ii = 2; ii <= nion; ++ii) ne = ne + (ii - 1) * nz[ii - 1];
        ne_sum = ne_sum + natomfr[iz - 1] * ne;
    }
    *nelec_new = ne_sum;
    return 0;
}

/* ionsmod.f:160-309 ions_solve_eos: secant iteration on nelec with the reference's bracket
 * bookkeeping; n receives the populations of the last Saha evaluation */
static int or_solve_eos(const or_eos_data *ed, const double *natomfr, double temp, double ndens,
                        double acc, double *nelec, double *n, int *niter)
{
    const double kti = 1.0 / (PC_KB * temp);
    const double sahac = pow(PC_PI2 * PC_ME * PC_KB * temp / (PC_H * PC_H), 1.5);
    double nel[2] = {0.0, 0.0}, err2[2] = {1e10, -1e10};
    int iprev = 0, iconv;
    for (iconv = 1; iconv <= 40; ++iconv) {
        const double sahac2 = 2.0 / (*nelec * ndens) * sahac;
        double nelec_new;
        if (or_saha_nelec(ed, natomfr, kti, sahac2, n, &nelec_new)) return 1;
        const double err = *nelec / nelec_new - 1.0;
        if (fabs(err) < acc) {
            if (iconv > 1) *nelec = nelec_new;
            break;
        }
        int ih = (err * err2[1] > err * err2[0]) ? 1 : 0;     /* maxloc(err*nec%err) */
        nel[ih] = *nelec;
        err2[ih] = err;
        if (iconv == 1) {
            double x = 1.5 * nelec_new - 0.5 * *nelec;
            x = fmin(3.0 * nelec_new, x);
            x = fmax(0.33 * nelec_new, x);
            *nelec = x;
        } else {
            ih = (fabs(err2[1]) > fabs(err2[0])) ? 2 : 1;      /* maxloc(abs(nec%err)) */
            const double help = fabs(err2[0] / err2[1]);
            double ynew;
            if (ih != iprev && help < 1e5 && help > 1e-5) ynew = 0.0;
            else ynew = 0.001 * err2[ih - 1];
            iprev = ih;
            const double dxdy = (nel[1] - nel[0]) / (err2[1] - err2[0]);
            *nelec = nel[0] - dxdy * (err2[0] - ynew);
            if (*nelec <= 0.0) return 1;
        }
    }
    *niter = iconv;
    return 0;
}

int or_eos(int nc, const or_eos_data *ed, const double *temp, const double *mass,
           const double *vol, const double *massfr, double acc, double *nelec, double *natom,
           double *natomfr, int nion, const int *ion_iz, const int *ion_ii, double *oc,
           double *ginv, int *niter)
{
    if (!(acc > 0.0)) acc = 1e-8;                  /* ionsmod.f:183 */
    int rc = 0;
    for (int i = 0; i < nc; ++i) {
        double *fr = natomfr + (size_t)i * 30;
        niter[i] = 0;
        if (mass[i] <= 0.0) continue;
        /* gas_setup.f:125-157 massfr2natomfr */
        const double *mf = massfr + (size_t)i * 30;
        double help = 0.0;
        for (int l = 0; l < 30; ++l) help = help + mf[l];
        double na = 0.0;
        for (int l = 0; l < 30; ++l) {
            fr[l] = (mf[l] / help) * mass[i];
            fr[l] = fr[l] / (ed->amass[l] * PC_AMU);
        }
        for (int l = 0; l < 30; ++l) na = na + fr[l];
        natom[i] = na;
        for (int l = 0; l < 30; ++l) fr[l] = fr[l] / na;
        /* eos_update.f:25-32 */
        const double ndens = natom[i] / vol[i];
        double n[30 * 31];
        memset(n, 0, sizeof n);
        if (or_solve_eos(ed, fr, temp[i], ndens, acc, &nelec[i], n, &niter[i])) { rc = 1; continue; }
        /* eos_update.f:36-45, for the listed ions */
        for (int k = 0; k < nion; ++k) {
            const int iz = ion_iz[k], ii = ion_ii[k];
            ginv[(size_t)i * nion + k] = 1.0 / EOS_G(iz, ii);
            oc[(size_t)i * nion + k] = natom[i] * fr[iz - 1] * EOS_G(iz, ii) * n[(iz - 1) * 31 + ii - 1] /
                                       (EOS_Q(iz, ii) * vol[i]);
        }
    }
    return rc;
}

/* ------------------------------------------------------------------------ */
/* SOURCE/sourceenergy.f90:23-50 (surface-like source) */
/* nltemod.f:231-260 */
void or_nlte_rad_est(int nc, int ng, const double *wl, const double *vol, const double *tally,
                     double *radtemp, double *jrad)
{
    for (int i = 0; i < nc; ++i) {
        double trad = pow(tally[i] / (PC_ACOEF * vol[i]), 0.25);
        radtemp[i] = radtemp[i] + 0.5 * (trad - radtemp[i]);
        if (!jrad) continue;
        for (int ig = 1; ig <= ng; ++ig) {
            double wlm = 0.5 * (wl[ig] + wl[ig - 1]);
            jrad[(size_t)i * ng + ig - 1] = wlm * wlm * or_planck(wlm, radtemp[i]) / PC_C;
        }
    }
}

void or_sourceenergy(int nc, const double *rcell, const double *radtemp, double lbol,
                     double *emit, double *sthermal)
{
    double s = 0.0;
    for (int i = 0; i < nc; ++i) {
        double t2 = radtemp[i] * radtemp[i];
        emit[i] = PC_PI * PC_C * (rcell[i] * rcell[i]) * PC_ACOEF * (t2 * t2);
        s += emit[i];
    }
    double s2 = 0.0;
    for (int i = 0; i < nc; ++i) { emit[i] = emit[i] / s * lbol; s2 += emit[i]; }
    *sthermal = s2;
}

/* SOURCE/sourcenumbers.f90:26-169 with one rank, in_srcepwr = 1, no surface source */
int64_t or_sourcenumbers(int nc, const double *emit, int64_t nstot, int *nvol)
{
    double etot = 0.0;
    int64_t ncactive = 0;
    for (int l = 0; l < nc; ++l) { etot += emit[l]; if (emit[l] > 0.0) ncactive++; }
    int64_t nsavail = nstot;
    double edone = 0.0;
    int64_t ndone = 0;
    double invn = 1.0 / (double)nsavail;
    double einv = etot == 0.0 ? 0.0 : 1.0 / etot;
    for (int l = 0; l < nc; ++l) {
        nvol[l] = 0;
        double en = emit[l];
        if (en == 0.0) continue;
        int64_t n = (int64_t)(en * (double)nsavail * einv);
        if (n < 1) n = 1;
        if (edone * einv > (double)ndone * invn) n = n + 1;
        nvol[l] = (int)n;
        edone += en;
        ndone += n;
    }
    int64_t sum = 0;
    for (int l = 0; l < nc; ++l) sum += nvol[l];
    int64_t nextra = nstot - sum;
    int64_t nsmean = nstot / ncactive;
    while (nextra != 0) {
        for (int l = 0; l < nc; ++l) {
            if (nextra == 0) break;
            if (nvol[l] < nsmean) continue;
            int sgn = nextra > 0 ? 1 : -1;
            nvol[l] += sgn;
            nextra -= sgn;
        }
    }
    sum = 0;
    for (int l = 0; l < nc; ++l) sum += nvol[l];
    return sum;
}

/* ------------------------------------------------------------------------ */
/* SOURCE/interior_source.f90:57-168 for the particles of one cell, then
 * SOURCE/source_transformdirection.f90:23-46 (direction2lab1, transportmod.f:129-145). */
static void or_sample_source(const or_model *m, const double *srcp, int i, int nemit,
                             or_stream *s, double *px, double *pmu, double *pom, double *pe,
                             double *pe0, double *pwl)
{
    const int ng = m->ng, nx = m->nx;
    double r1 = or_draw(s);
    double denom2 = 0.0;
    int iig = 1;
    for (int ig = 1; ig <= ng; ++ig) {
        iig = ig;
        if (r1 >= denom2 && r1 < denom2 + srcp[ig - 1]) break;
        denom2 = denom2 + srcp[ig - 1];
    }
    r1 = or_draw(s);
    double wl0 = 1.0 / ((1.0 - r1) * WLINV(iig) + r1 * WLINV(iig + 1));
    double ep0 = m->emit[i - 1] / (double)nemit;
    r1 = or_draw(s);
    double mu0 = sqrt(r1);
    r1 = or_draw(s);
    double om0 = PC_PI2 * r1;
    int l1 = i - 1 < 1 ? 1 : i - 1, l2 = i + 1 > nx ? nx : i + 1;
    double uul = 0.5 * (m->emit[l1 - 1] + m->emit[i - 1]);
    double uur = 0.5 * (m->emit[l2 - 1] + m->emit[i - 1]);
    double uumax = uul > uur ? uul : uur;
    uul = uul / uumax;
    uur = uur / uumax;
    double r2 = 1.0, r3 = 0.0, x0 = XA(i);
    const double third = (double)(1.0f / 3.0f);       /* single-precision 1.0/3.0 */
    double xl3 = XA(i) * XA(i) * XA(i), xr3 = XA(i + 1) * XA(i + 1) * XA(i + 1);
    while (r2 > r3) {
        r1 = or_draw(s);
        x0 = pow(r1 * xr3 + (1.0 - r1) * xl3, third);
        r3 = (x0 - XA(i)) / DX(i);
        r3 = r3 * uur + (1.0 - r3) * uul;
        r2 = or_draw(s);
    }
    /* source_transformdirection */
    int ix = or_binsrch(x0, m->xarr, nx + 1, 0);
    double h1 = XA(ix), h2 = XA(ix + 1);
    double vx = (VXA(ix) * (h2 - x0) + VXA(ix + 1) * (x0 - h1)) / (h2 - h1);
    double cmffact = 1.0 + mu0 * vx * CINV;
    double mu = (mu0 + vx * CINV) / cmffact;
    if (mu > 1.0) mu = 1.0;
    if (mu < -1.0) mu = -1.0;
    *px = x0; *pmu = mu; *pom = om0; *pe = ep0; *pe0 = ep0; *pwl = wl0;
}

static double *or_source_probs(const or_model *m)
{
    const int ng = m->ng;
    double *p = malloc(sizeof(double) * (size_t)ng * m->nx);
    for (int l = 1; l <= m->nx; ++l) {
        double *pl = p + (size_t)(l - 1) * ng;
        or_specintv(1.0 / m->radtemp[l - 1], m->wlinv, ng, 1, pl);
        double s = 0.0;
        for (int g = 0; g < ng; ++g) s += pl[g];
        for (int g = 0; g < ng; ++g) pl[g] = pl[g] / s;
    }
    return p;
}

int or_sources(const or_model *m, uint32_t seed, uint32_t iteration,
               double *x, double *mu, double *om, double *e, double *e0, double *wl)
{
    double *srcp = or_source_probs(m);
    int64_t p = 0;
    for (int i = 1; i <= m->nx; ++i) {
        int n = m->nvol[i - 1];
        for (int k = 0; k < n; ++k, ++p) {
            or_stream s;
            or_stream_init(&s, seed, iteration, (uint32_t)p, 0u);
            or_sample_source(m, srcp + (size_t)(i - 1) * m->ng, i, n, &s,
                             &x[p], &mu[p], &om[p], &e[p], &e0[p], &wl[p]);
        }
    }
    free(srcp);
    return 0;
}

/* ------------------------------------------------------------------------ */
typedef struct {
    double x, mu, e, e0, wl;
    int ix, ic, ig, itype, idist, stat;   /* stat 0 live, 1 flux, 2 dead */
    int nnewton;                           /* Newton-Raphson Doppler distances taken */
} or_ptcl;

typedef struct {
    double *specarr;    /* ng, specintv(tempinv(ic)) for the cached cell */
    int *glumps;        /* ng */
    unsigned char *llumps;
    int ic, nlump;
    double capgreyinv, capemitgreyinv, speclump, emitlump, caplump, doplump;
} or_cache;

/* diffusion11.f90:95-152: cell quantities (recomputed whenever the cell changes) */
static void or_cache_cell(const or_model *m, or_cache *c, int ic, int ix)
{
    const int ng = m->ng;
    double dist = DX(ix);
    c->ic = ic;
    or_specintv(TINV(ic), m->wlinv, ng, 1, c->specarr);
    c->capgreyinv = 1.0 / m->capgrey[ic - 1];
    if (!(c->capgreyinv >= 0.0)) c->capgreyinv = 0.0;
    c->capemitgreyinv = 0.0;
    if (NLTE) {
        c->capemitgreyinv = 1.0 / m->capemitgrey[ic - 1];
        if (!(c->capemitgreyinv >= 0.0)) c->capemitgreyinv = 0.0;
    }
    int glump = 0, gunlump = ng;
    for (int k = 0; k < ng; ++k) c->glumps[k] = 0;
    c->speclump = OPL(7, ic);
    if (c->speclump != 0.0) {
        for (int iig = 1; iig <= ng; ++iig) {
            if (CAP(iig, ic) * dist >= m->taulump && (SIG(ic) + CAP(iig, ic)) * dist >= m->tauddmc) {
                c->llumps[iig - 1] = 1;
                glump++;
                c->glumps[glump - 1] = iig;
            } else {
                c->llumps[iig - 1] = 0;
                c->glumps[gunlump - 1] = iig;
                gunlump--;
            }
        }
    }
    if (glump == ng) {
        c->emitlump = 1.0;
        c->caplump = m->capgrey[ic - 1];
        c->doplump = 0.0;
    } else {
        c->caplump = OPL(8, ic) * c->speclump;
        double el = NLTE ? OPL(9, ic) * c->capemitgreyinv : OPL(8, ic) * c->capgreyinv;
        c->emitlump = el < 1.0 ? el : 1.0;
        c->doplump = OPL(11, ic) * c->speclump;
    }
    c->nlump = glump;
}

static inline void vinterp(const or_model *m, int ix, double x, double *vx)
{
    double h1 = XA(ix), h2 = XA(ix + 1);
    *vx = (VXA(ix) * (h2 - x) + VXA(ix + 1) * (x - h1)) / DX(ix);
}

/* transport11.f90:259-285: the experimental anisotropic Thomson cosine.  The second cube
 * root has a negative base for r1 below ~0.957, where the real power is NaN; every range
 * test then fails and the isotropic cosine is taken, as in the reference. */
static double or_thomson_aniso_mu(double r1)
{
    const double q = sqrt(1.0 + 16.0 * ((r1 - 0.5) * (r1 - 0.5)));
    const double mutemp = pow(4.0 * r1 - 2.0 + q, 1.0 / 3.0) - pow(q - (4.0 - 2.0 * r1), 1.0 / 3.0);
    const double muhelp = mutemp;
    const double B = 2 * mutemp * muhelp;
    const double C = mutemp * mutemp + muhelp * muhelp - 1;
    const double munew1 = (-B + sqrt(B * B - 4 * C)) / 2;
    const double munew2 = (-B - sqrt(B * B - 4 * C)) / 2;
    const int ok1 = munew1 <= 1 && munew1 >= -1, ok2 = munew2 <= 1 && munew2 >= -1;
    if (ok1 && ok2) return r1 < 0.5 ? munew1 : munew2;
    if (ok1) return munew1;
    if (ok2) return munew2;
    return 1.0 - 2.0 * r1;
}

/* TRANSPORT1/transport11.f90:86-489 — one IMC event */
static int or_transport11(const or_model *m, or_ptcl *p, double *vx, or_stream *s,
                          double *eraddens, double *jrad, double *eamp, double *totevelo)
{
    const int ng = m->ng, nx = m->nx;
    int ix = p->ix, ic = p->ic, ig = p->ig;
    double x = p->x, mu = p->mu, e = p->e, e0 = p->e0, wl = p->wl;
    double r1, r2, db, dcol, dthm, ddop, d;
    int ixnext;
    *eraddens = 0.0; *jrad = 0.0; *eamp = 0.0;
    double elabfact = 1.0 - mu * (*vx) * CINV;
    double elabfactold = elabfact;
    double far = 2.0 * fabs(sqrt(XA(ix + 1) * XA(ix + 1) - XA(ix) * XA(ix)));
    /* shell distance :95-103 */
    if (ix == 1 || mu >= -sqrt(1.0 - (XA(ix) / x) * (XA(ix) / x))) {
        db = fabs(sqrt(XA(ix + 1) * XA(ix + 1) - (1.0 - mu * mu) * x * x) - mu * x);
        ixnext = ix + 1;
    } else {
        db = fabs(sqrt(XA(ix) * XA(ix) - (1.0 - mu * mu) * x * x) + mu * x);
        ixnext = ix - 1;
    }
    if (db != db) return 1;
    if (SIG(ic) > 0.0) {
        r1 = or_draw(s);
        dthm = -log(r1) / (elabfact * SIG(ic));
    } else dthm = far;
    if (CAP(ig, ic) <= 0.0) dcol = far;
    else {
        r1 = or_draw(s);
        dcol = -log(r1) / (elabfact * CAP(ig, ic));
    }
    /* Doppler distance :129-199 */
    double dvdr = (VXA(ix + 1) - VXA(ix)) / DX(ix);
    double vhelp = VXA(ix) - dvdr * XA(ix);
    if (fabs(*vx - vhelp) / PC_C >= 0.1) {
        p->nnewton++;
        const double epstol = 1e-3;
        int it = 0;
        double shelp = db;
        double xhelp = sqrt(x * x + shelp * shelp + 2.0 * shelp * x * mu);
        double muhelp = (mu * x + shelp) / xhelp;
        double phi = dvdr * mu * x - PC_C * (1 - wl * WLINV(ig + 1)) + vhelp * muhelp + dvdr * shelp;
        double dphi = dvdr + vhelp * (1 - muhelp * muhelp) / xhelp;
        ddop = shelp - phi / dphi;
        while ((fabs(ddop - shelp) > ddop * epstol) && it < 100) {
            it = it + 1;
            shelp = ddop;
            xhelp = sqrt(x * x + shelp * shelp + 2.0 * shelp * x * mu);
            muhelp = (mu * x + shelp) / xhelp;
            if (dvdr > 0 && ig < ng) {
                phi = dvdr * mu * x - PC_C * (1 - wl * WLINV(ig + 1)) + vhelp * muhelp + dvdr * shelp;
            } else if (dvdr < 0 && ig > 1) {
                phi = dvdr * mu * x - PC_C * (1 - wl * WLINV(ig)) + vhelp * muhelp + dvdr * shelp;
            } else {
                ddop = far;
                break;
            }
            dphi = dvdr + vhelp * (1 - muhelp * muhelp) / xhelp;
            ddop = shelp - phi / dphi;
        }
        if (ddop <= 0.0 || it == 100) ddop = far;
    } else {
        if (dvdr > 0 && ig < ng) ddop = 3.0 * PC_C / DIVV(ix) * (elabfact - wl * WLINV(ig + 1));
        else if (dvdr < 0 && ig > 1) ddop = 3.0 * PC_C / DIVV(ix) * (elabfact - wl * WLINV(ig));
        else ddop = far;
        if (ddop < 0.0) ddop = far;
    }
    /* minimum distance :204-214 (minloc picks the first minimum) */
    double darr[4] = {dcol, dthm, ddop, db};
    int idist = 1;
    d = darr[0];
    for (int k = 1; k < 4; ++k) if (darr[k] < d) { d = darr[k]; idist = k + 1; }
    p->idist = idist;
    for (int k = 0; k < 4; ++k) if (darr[k] != darr[k]) return 3;
    if (d < 0.0) return 4;
    /* move :217-230 */
    double xold = x;
    x = sqrt(x * x + d * d + 2.0 * d * x * mu);
    {
        double h1 = XA(ix), h2 = XA(ix + 1);
        *vx = (VXA(ix) * (h2 - x) + VXA(ix + 1) * (x - h1)) / DX(ix);
    }
    if (x == 0.0) mu = 1.0;
    else mu = (xold * mu + d) / x;
    /* tally :233-236 */
    *eraddens = e * elabfact * elabfact * d * CINV;
    *jrad = PC_C * (*eraddens);
    elabfact = 1.0 - mu * (*vx) * CINV;
    /* inner hole :243-252 */
    if (x < XA(1)) {
        x = XA(1);
        *vx = VXA(1);
        ix = 1;
        ic = 1;
        r1 = or_draw(s);
        mu = 1.0 - r1;
        goto done;
    }
    /* collisions :256-296 */
    if (d == dthm || d == dcol) {
        r1 = or_draw(s);
        if (d == dthm && m->thm_aniso) mu = or_thomson_aniso_mu(r1);
        else mu = 1.0 - 2.0 * r1;
        mu = (mu + (*vx) * CINV) / (1.0 + (*vx) * mu * CINV);
    } else if (d == db) {
        if (mu >= 0.0 && ix == nx) {
            p->stat = 1;
            goto done;
        }
    }
    if (d == dthm) {
        wl = wl * (1.0 - mu * (*vx) * CINV) / elabfact;
        double help = elabfact / (1.0 - mu * (*vx) * CINV);
        *totevelo = *totevelo + e * (1.0 - help);
        e = e * help;
        e0 = e0 * help;
    } else if (d == dcol) {
        r1 = or_draw(s);                       /* unused draw, :313 */
        double help = elabfact / (1.0 - mu * (*vx) * CINV);
        *totevelo = *totevelo + e * (1.0 - help);
        e = e * help;
        e0 = e0 * help;
        if (ng != 1) {
            r1 = or_draw(s);
            ig = or_emitgroup(m, r1, ic);
            if (((SIG(ic) + CAP(ig, ic)) * DX(ix) >= m->tauddmc) && !m->puretran) p->itype = 2;
        }
        if (p->itype == 2) {
            *totevelo = *totevelo + e * (*vx) * mu * CINV;
            e = e * (1.0 - (*vx) * mu * CINV);
            e0 = e0 * (1.0 - (*vx) * mu * CINV);
        } else {
            r1 = or_draw(s);
            wl = 1.0 / ((1.0 - r1) * WLINV(ig) + r1 * WLINV(ig + 1));
            wl = wl * (1.0 - (*vx) * mu * CINV);
        }
    } else if (d == db && ixnext > ix && ix == nx) {
        /* outer boundary reached with a rounding-negative cosine: escapes (the reference
         * would index the non-existent cell nx+1) */
        p->stat = 1;
        goto done;
    } else if (d == db && ixnext > ix) {
        int l = ix + 1;
        if ((SIG(l) + CAP(ig, l)) * DX(ix + 1) < m->tauddmc || m->puretran) {
            x = XA(ix + 1);
            *vx = VXA(ix + 1);
            ix = ix + 1;
            ic = ix;
        } else {
            mu = (mu - (*vx) * CINV) / (1.0 - (*vx) * mu * CINV);
            double help = (CAP(ig, l) + SIG(l)) * DX(ix + 1);
            help = 4.0 / (3.0 * help + 6.0 * PC_DEXT);
            r1 = or_draw(s);
            if (r1 < help * (1.0 + 1.5 * fabs(mu))) {
                p->itype = 2;
                *totevelo = *totevelo + e * (1.0 - elabfact);
                e = e * elabfact;
                e0 = e0 * elabfact;
                wl = wl / elabfact;
                ix = ix + 1;
                ic = ix;
            } else {
                r1 = or_draw(s);
                r2 = or_draw(s);
                mu = -(r1 > r2 ? r1 : r2);
                mu = (mu + (*vx) * CINV) / (1.0 + (*vx) * mu * CINV);
                x = XA(ix + 1);
                *vx = VXA(ix + 1);
            }
        }
    } else if (d == db) {
        int l = ix - 1;
        if ((SIG(l) + CAP(ig, l)) * DX(ix - 1) < m->tauddmc || m->puretran) {
            x = XA(ix);
            *vx = VXA(ix);
            ix = ix - 1;
            ic = ix;
        } else {
            mu = (mu - (*vx) * CINV) / (1.0 - (*vx) * mu * CINV);
            if (m->ampfact && mu < 0.0) {
                /* amplification factor :410-421 */
                double h = 1.0 / fabs(mu);
                h = h < 100.0 ? h : 100.0;
                *totevelo = *totevelo - e * 2.0 * (0.55 * h - 1.25 * fabs(mu)) * (*vx) * CINV;
                *eamp = e * 2.0 * 0.55 * (h - 2.0 > 0.0 ? h - 2.0 : 0.0) * (*vx) * CINV;
                h = h < 2.0 ? h : 2.0;
                e0 = e0 * (1.0 + 2.0 * (0.55 * h - 1.25 * fabs(mu)) * (*vx) * CINV);
                e = e * (1.0 + 2.0 * (0.55 * h - 1.25 * fabs(mu)) * (*vx) * CINV);
            }
            double help = (CAP(ig, l) + SIG(l)) * DX(ix - 1);
            help = 4.0 / (3.0 * help + 6.0 * PC_DEXT);
            r1 = or_draw(s);
            if (r1 < help * (1.0 + 1.5 * fabs(mu))) {
                p->itype = 2;
                *totevelo = *totevelo + e * (1.0 - elabfact);
                e = e * elabfact;
                e0 = e0 * elabfact;
                wl = wl / elabfact;
                ix = ix - 1;
                ic = ix;
            } else {
                r1 = or_draw(s);
                r2 = or_draw(s);
                mu = r1 > r2 ? r1 : r2;
                mu = (mu + (*vx) * CINV) / (1.0 + (*vx) * mu * CINV);
                x = XA(ix);
                *vx = VXA(ix);
            }
        }
    } else if (d == ddop) {
        if (dvdr > 0) {
            if (ig < ng) {
                ig = ig + 1;
                wl = (WL(ig) + 1e-6 * (WL(ig + 1) - WL(ig))) * elabfactold;
            } else {
                r1 = or_draw(s);
                wl = 1.0 / (r1 * WLINV(ng + 1) + (1.0 - r1) * WLINV(ng));
                wl = wl * elabfact;
            }
        } else if (dvdr < 0) {
            if (ig > 1) {
                wl = (WL(ig) + 1e-6 * (WL(ig - 1) - WL(ig))) * elabfactold;
                ig = ig - 1;
            } else {
                r1 = or_draw(s);
                wl = 1.0 / (r1 * WLINV(1) + (1.0 - r1) * WLINV(2));
                wl = wl * elabfact;
            }
        }
        if (((SIG(ic) + CAP(ig, ic)) * DX(ix) >= m->tauddmc) && !m->puretran) {
            p->itype = 2;
            *totevelo = *totevelo + e * (*vx) * mu * CINV;
            e = e * elabfact;
            e0 = e0 * elabfact;
            wl = wl / elabfact;
        }
    } else {
        p->x = x; p->mu = mu; p->e = e; p->e0 = e0; p->wl = wl;
        p->ix = ix; p->ic = ic; p->ig = ig;
        return 17;
    }
done:
    p->x = x; p->mu = mu; p->e = e; p->e0 = e0; p->wl = wl;
    p->ix = ix; p->ic = ic; p->ig = ig;
    return 0;
}

/* resolved leakage opacity of group g from cell ic through face `side`
 * (diffusion11.f90:383-395 left, :496-508 right, :455-457 outer boundary) */
static inline double or_resleak(const or_model *m, int ix, int ic, int g, int side)
{
    if (side < 0) {
        int l = ix - 1;
        if ((CAP(g, l) + SIG(l)) * DX(ix - 1) < m->tauddmc) {
            double mfp = (CAP(g, ic) + SIG(ic)) * DX(ix);
            double ppl = 4.0 / (3.0 * mfp + 6.0 * PC_DEXT);
            return 1.5 * ppl * (XA(ix) * XA(ix)) / DX3(ix);
        }
        double mfp = (SIG(ic) + CAP(g, ic)) * DX(ix) + (SIG(l) + CAP(g, l)) * DX(ix - 1);
        return 2.0 * (XA(ix) * XA(ix)) / (mfp * DX3(ix));
    }
    if (ix == m->nx) {
        double mfp = (CAP(g, ic) + SIG(ic)) * DX(ix);
        double ppr = 4.0 / (3.0 * mfp + 6.0 * PC_DEXT);
        return 1.5 * ppr * (XA(ix + 1) * XA(ix + 1)) / DX3(ix);
    }
    int l = ix + 1;
    if ((CAP(g, l) + SIG(l)) * DX(ix + 1) < m->tauddmc) {
        double mfp = (CAP(g, ic) + SIG(ic)) * DX(ix);
        double ppr = 4.0 / (3.0 * mfp + 6.0 * PC_DEXT);
        return 1.5 * ppr * (XA(ix + 1) * XA(ix + 1)) / DX3(ix);
    }
    double mfp = (SIG(ic) + CAP(g, ic)) * DX(ix) + (SIG(l) + CAP(g, l)) * DX(ix + 1);
    return 2.0 * (XA(ix + 1) * XA(ix + 1)) / (mfp * DX3(ix));
}

/* TRANSPORT1/diffusion11.f90:91-631 — one DDMC event, perfect cell cache */
static int or_diffusion11(const or_model *m, or_ptcl *p, double *vx, or_cache *c, or_stream *s,
                          double *eraddens, double *jrad, double *totevelo)
{
    const int ng = m->ng, nx = m->nx;
    int ix = p->ix, ic = p->ic, ig = p->ig;
    double x = p->x, mu = p->mu, e = p->e, e0 = p->e0, wl = p->wl;
    double r1, r2;
    int iiig = ig;
    *eraddens = 0.0; *jrad = 0.0;
    double dist = DX(ix);
    if (ic != c->ic) or_cache_cell(m, c, ic, ix);
    int glump = (CAP(ig, ic) * dist >= m->taulump) ? c->nlump : 0;
    if ((SIG(ic) + CAP(ig, ic)) * dist < m->tauddmc) return 100;
    double emitlump, caplump, doplump;
    if (glump > 0) {
        emitlump = c->emitlump;
        caplump = c->caplump;
        doplump = c->doplump;
    } else {
        emitlump = NLTE ? or_specint0(TINV(ic), m->wlinv, ig) * c->capemitgreyinv * CAPE(ig, ic)
                        : or_specint0(TINV(ic), m->wlinv, ig) * c->capgreyinv * CAP(ig, ic);
        caplump = CAP(ig, ic);
        if (ig != ng && DIVV(ix) >= 0)
            doplump = or_dopspeccalc(TINV(ic), m->wlinv, ig) * DIVV(ix) / (c->specarr[ig - 1] * 3 * PC_C);
        else if (ig != 1 && DIVV(ix) < 0)
            doplump = -1 * or_dopspeccalc(TINV(ic), m->wlinv, ig - 1) * DIVV(ix) / (c->specarr[ig - 1] * 3 * PC_C);
        else
            doplump = 0.0;
    }
    double opacleak[2];
    if (glump > 0) {
        opacleak[0] = OPL(1, ic);
        opacleak[1] = OPL(2, ic);
    } else {
        opacleak[0] = (ix == 1) ? 0.0 : or_resleak(m, ix, ic, ig, -1);
        opacleak[1] = or_resleak(m, ix, ic, ig, +1);
    }
    double denom = opacleak[0] + opacleak[1] + (1.0 - emitlump) * caplump + doplump;
    denom = 1.0 / denom;
    r1 = or_draw(s);
    double tau = fabs(log(r1) * denom * CINV);
    double ddmct = tau;
    *eraddens = e * ddmct;
    *jrad = PC_C * (*eraddens);
    r1 = or_draw(s);
    double probleak[2] = {opacleak[0] * denom, opacleak[1] * denom};
    const double pa = 0.0;
    double pdop = doplump * denom;

    if (r1 >= pa && r1 < pa + pdop) {
        /* Doppler shift :283-355 */
        if (glump == 0) iiig = ig;
        else {
            r1 = or_draw(s);
            double denom2 = 0.0, help = 1.0 / doplump;
            if (DIVV(ix) >= 0) {
                for (int iig = 1; iig <= glump; ++iig) {
                    iiig = c->glumps[iig - 1];
                    if (iiig == ng) continue;
                    if (CAP(iiig + 1, ic) * dist >= m->taulump) continue;
                    double res = or_dopspeccalc(TINV(ic), m->wlinv, iiig) * DIVV(ix) / (3 * PC_C);
                    denom2 = denom2 + res * c->speclump * help;
                    if (denom2 > r1) break;
                }
            } else {
                for (int iig = glump; iig >= 1; --iig) {
                    iiig = c->glumps[iig - 1];
                    if (iiig == 1) continue;
                    if (CAP(iiig - 1, ic) * dist >= m->taulump) continue;
                    double res = or_dopspeccalc(TINV(ic), m->wlinv, iiig - 1) * DIVV(ix) / (3 * PC_C);
                    denom2 = denom2 - res * c->speclump * help;
                    if (denom2 > r1) break;
                }
            }
        }
        if (DIVV(ix) >= 0) {
            ig = iiig + 1;
            wl = WL(ig);
            if (ig > ng) ig = ng;
        } else {
            ig = iiig - 1;
            wl = WL(ig + 1);
            if (ig < 1) ig = 1;
        }
        if ((SIG(ic) + CAP(ig, ic)) * dist < m->tauddmc) {
            p->itype = 1;
            r1 = or_draw(s);
            mu = 1.0 - 2.0 * r1;
            r1 = or_draw(s);
            const double third = (double)(1.0f / 3.0f);
            x = pow(r1 * (XA(ix + 1) * XA(ix + 1) * XA(ix + 1)) + (1.0 - r1) * (XA(ix) * XA(ix) * XA(ix)), third);
            if (x > XA(ix + 1)) x = XA(ix + 1);
            if (x < XA(ix)) x = XA(ix);
            vinterp(m, ix, x, vx);
            mu = (mu + (*vx) * CINV) / (1.0 + (*vx) * mu * CINV);
            wl = wl * (1.0 - (*vx) * mu * CINV);
            double help = 1.0 / (1.0 - (*vx) * mu * CINV);
            *totevelo = *totevelo + e * (1.0 - help);
            e = e * help;
            e0 = e0 * help;
        }
    } else if (r1 >= pa + pdop && r1 < pa + pdop + probleak[0]) {
        /* left leakage :358-430 */
        p->idist = -3;
        if (ix == 1) {
            goto ret;   /* returns before the adiabatic loss */
        } else {
            int l = ix - 1;
            if (glump == 0) iiig = ig;
            else {
                r1 = or_draw(s);
                double denom2 = 0.0, help = 1.0 / opacleak[0];
                for (int iig = 1; iig <= glump; ++iig) {
                    iiig = c->glumps[iig - 1];
                    double specig = c->specarr[iiig - 1];
                    double res = or_resleak(m, ix, ic, iiig, -1);
                    denom2 = denom2 + specig * res * c->speclump * help;
                    if (denom2 > r1) break;
                }
            }
            if ((SIG(l) + CAP(iiig, l)) * DX(ix - 1) < m->tauddmc) {
                p->itype = 1;
                r1 = or_draw(s);
                wl = 1.0 / (r1 * WLINV(iiig + 1) + (1.0 - r1) * WLINV(iiig));
                x = XA(ix);
                *vx = VXA(ix);
                r1 = or_draw(s);
                r2 = or_draw(s);
                mu = -(r1 > r2 ? r1 : r2);
                mu = (mu + (*vx) * CINV) / (1.0 + (*vx) * mu * CINV);
                double help = 1.0 / (1.0 - (*vx) * mu * CINV);
                *totevelo = *totevelo + e * (1.0 - help);
                e = e * help;
                e0 = e0 * help;
                wl = wl * (1.0 - (*vx) * mu * CINV);
            }
            ix = ix - 1;
            ic = ix;
            ig = iiig;
        }
    } else if (r1 >= pa + pdop + probleak[0] && r1 < pa + pdop + (probleak[0] + probleak[1])) {
        /* right leakage :434-545 */
        p->idist = -4;
        if (ix == nx) {
            p->stat = 1;
            r1 = or_draw(s);
            r2 = or_draw(s);
            mu = r1 > r2 ? r1 : r2;
            if (glump != 0) {
                r1 = or_draw(s);
                double denom2 = 0.0, help = 1.0 / opacleak[1];
                for (int iig = 1; iig <= glump; ++iig) {
                    iiig = c->glumps[iig - 1];
                    double specig = c->specarr[iiig - 1];
                    double res = or_resleak(m, ix, ic, iiig, +1);
                    denom2 = denom2 + specig * res * c->speclump * help;
                    if (denom2 > r1) break;
                }
                ig = iiig;
            }
            r1 = or_draw(s);
            wl = 1.0 / (r1 * WLINV(ig + 1) + (1.0 - r1) * WLINV(ig));
            x = XA(nx + 1);
            *vx = VXA(nx + 1);
            mu = (mu + (*vx) * CINV) / (1.0 + (*vx) * mu * CINV);
            if (mu > 1.0) mu = 1.0;
            double help = 1.0 / (1.0 - mu * (*vx) * CINV);
            *totevelo = *totevelo + e * (1.0 - help);
            wl = wl / help;
            e = e * help;
            e0 = e0 * help;
            goto ret;
        } else {
            int l = ix + 1;
            if (glump == 0) iiig = ig;
            else {
                r1 = or_draw(s);
                double denom2 = 0.0, help = 1.0 / opacleak[1];
                for (int iig = 1; iig <= glump; ++iig) {
                    iiig = c->glumps[iig - 1];
                    double specig = c->specarr[iiig - 1];
                    double res = or_resleak(m, ix, ic, iiig, +1);
                    denom2 = denom2 + specig * res * c->speclump * help;
                    if (denom2 > r1) break;
                }
            }
            if ((SIG(l) + CAP(iiig, l)) * DX(ix + 1) < m->tauddmc) {
                p->itype = 1;
                r1 = or_draw(s);
                wl = 1.0 / (r1 * WLINV(iiig + 1) + (1.0 - r1) * WLINV(iiig));
                x = XA(ix + 1);
                *vx = VXA(ix + 1);
                r1 = or_draw(s);
                r2 = or_draw(s);
                mu = r1 > r2 ? r1 : r2;
                mu = (mu + (*vx) * CINV) / (1.0 + (*vx) * mu * CINV);
                double help = 1.0 / (1.0 - (*vx) * mu * CINV);
                *totevelo = *totevelo + e * (1.0 - help);
                e = e * help;
                e0 = e0 * help;
                wl = wl * (1.0 - (*vx) * mu * CINV);
            }
            ix = ix + 1;
            ic = ix;
            ig = iiig;
        }
    } else {
        /* effective scattering :549-620 */
        p->idist = -2;
        if (glump == ng) return 102;
        if (glump == 0) {
            r1 = or_draw(s);
            iiig = or_emitgroup(m, r1, ic);
        } else {
            r1 = or_draw(s);
            double denom2 = 1.0 / (1.0 - emitlump);
            double denom3 = 0.0;
            for (int iig = ng; iig >= glump + 1; --iig) {
                iiig = c->glumps[iig - 1];
                double help = NLTE ? c->specarr[iiig - 1] * CAPE(iiig, ic) * c->capemitgreyinv
                                   : c->specarr[iiig - 1] * CAP(iiig, ic) * c->capgreyinv;
                denom3 = denom3 + help * denom2;
                if (denom3 > r1) break;
            }
        }
        ig = iiig;
        if ((SIG(ic) + CAP(ig, ic)) * dist < m->tauddmc) {
            p->itype = 1;
            r1 = or_draw(s);
            wl = 1.0 / ((1.0 - r1) * WLINV(ig) + r1 * WLINV(ig + 1));
            r1 = or_draw(s);
            mu = 1.0 - 2.0 * r1;
            r1 = or_draw(s);
            const double third = (double)(1.0f / 3.0f);
            x = pow(r1 * (XA(ix + 1) * XA(ix + 1) * XA(ix + 1)) + (1.0 - r1) * (XA(ix) * XA(ix) * XA(ix)), third);
            if (x > XA(ix + 1)) x = XA(ix + 1);
            if (x < XA(ix)) x = XA(ix);
            vinterp(m, ix, x, vx);
            mu = (mu + (*vx) * CINV) / (1.0 + (*vx) * mu * CINV);
            double help = 1.0 / (1.0 - (*vx) * mu * CINV);
            *totevelo = *totevelo + e * (1.0 - help);
            e = e * help;
            e0 = e0 * help;
            wl = wl * (1.0 - (*vx) * mu * CINV);
        }
    }
    if (x < XA(1)) {
        p->stat = 2;
        goto ret;
    }
    {
        double f = exp(-DIVV(ix) * ddmct / 3);
        *totevelo = *totevelo + e * (1.0 - f);
        e = e * f;
        e0 = e0 * f;
    }
ret:
    p->x = x; p->mu = mu; p->e = e; p->e0 = e0; p->wl = wl;
    p->ix = ix; p->ic = ic; p->ig = ig;
    return 0;
}

/* particle_advance.f90:140-294 for one particle, then fluxtally.f:30-54 */
int or_iteration(const or_model *m, uint32_t seed, uint32_t iteration, or_result *res)
{
    const int ng = m->ng, nc = m->nx;
    int64_t npart = 0;
    for (int i = 0; i < nc; ++i) npart += m->nvol[i];
    int64_t *pstart = malloc(sizeof(int64_t) * (nc + 1));
    pstart[0] = 0;
    for (int i = 0; i < nc; ++i) pstart[i + 1] = pstart[i] + m->nvol[i];
    double *srcp = or_source_probs(m);

    memset(res->tally, 0, sizeof(double) * nc);
    memset(res->jrad, 0, sizeof(double) * (size_t)ng * nc);
    memset(res->luminos, 0, sizeof(double) * ng);
    memset(res->lumdev, 0, sizeof(double) * ng);
    memset(res->lumnum, 0, sizeof(int) * ng);
    if (res->eamp) memset(res->eamp, 0, sizeof(double) * nc);
    for (int k = 0; k < 7; ++k) res->tot[k] = 0.0;
    for (int k = 0; k < 10; ++k) res->ct[k] = 0;
    int any_err = 0;

#pragma omp parallel
    {
        double *tally = calloc(nc, sizeof(double));
        double *eamp = calloc(nc, sizeof(double));
        double *jrad = calloc((size_t)ng * nc, sizeof(double));
        double *lum = calloc(ng, sizeof(double)), *dev = calloc(ng, sizeof(double));
        int64_t *num = calloc(ng, sizeof(int64_t));
        double tot[7] = {0, 0, 0, 0, 0, 0, 0};
        int64_t ct[10] = {0};
        or_cache cache;
        memset(&cache, 0, sizeof cache);
        cache.specarr = malloc(sizeof(double) * ng);
        cache.glumps = malloc(sizeof(int) * ng);
        cache.llumps = malloc(ng);
        cache.ic = 0;
        int lerr = 0;
#pragma omp for schedule(dynamic, 64)
        for (int64_t ip = 0; ip < npart; ++ip) {
            int cell = 1;
            {   /* owning cell of particle ip (particles are numbered cell by cell) */
                int lo = 0, hi = nc;
                while (hi - lo > 1) { int mid = (lo + hi) / 2; if (pstart[mid] <= ip) lo = mid; else hi = mid; }
                cell = lo + 1;
            }
            or_stream ss;
            or_stream_init(&ss, seed, iteration, (uint32_t)ip, 0u);
            or_ptcl p;
            double om, x0, mu0, e0v, ev, wl0;
            or_sample_source(m, srcp + (size_t)(cell - 1) * ng, cell, m->nvol[cell - 1], &ss,
                             &x0, &mu0, &om, &ev, &e0v, &wl0);
            tot[5] += ev;
            or_stream st;
            or_stream_init(&st, seed, iteration, (uint32_t)ip, 1u);
            p.x = x0; p.mu = mu0; p.e = ev; p.e0 = e0v; p.wl = wl0;
            p.ix = or_binsrch(p.x, m->xarr, m->nx + 1, 0);
            p.ic = p.ix;
            p.ig = or_binsrch(p.wl, m->wl, ng + 1, 0);
            double help = DX(p.ix);
            double tau = (SIG(p.ic) + CAP(p.ig, p.ic)) * help;
            p.itype = (m->puretran || tau < m->tauddmc) ? 1 : 2;
            double vx;
            {
                double h1 = XA(p.ix), h2 = XA(p.ix + 1);
                vx = (VXA(p.ix) * (h2 - p.x) + VXA(p.ix + 1) * (p.x - h1)) / (h2 - h1);
            }
            if (p.itype == 1) {
                double labfact = 1.0 - vx * p.mu / PC_C;
                p.wl = p.wl * labfact;
                p.e = p.e / labfact;
                p.e0 = p.e0 / labfact;
            }
            tot[3] += p.e;
            ct[0]++;
            p.stat = 0;
            p.idist = 0;
            p.nnewton = 0;
            int64_t istep = 0;
            int perr = 0;
            while (p.stat == 0) {
                istep++;
                or_stream_step(&st, (uint32_t)istep);
                int icold = p.ic, igold = p.ig;
                double erd, jr;
                int ierr;
                if (p.itype == 1 || m->puretran) {
                    ct[1]++;
                    double ea;
                    ierr = or_transport11(m, &p, &vx, &st, &erd, &jr, &ea, &tot[1]);
                    if (m->ampfact) {                 /* particle_advance.f90:216-217 */
                        eamp[icold - 1] += ea;
                        tot[6] += ea;
                    }
                    if (p.itype != 1) ct[4]++;
                } else {
                    ct[2]++;
                    ierr = or_diffusion11(m, &p, &vx, &cache, &st, &erd, &jr, &tot[1]);
                    if (p.itype == 1) ct[4]++;
                }
                tally[icold - 1] += erd;
                jrad[(size_t)(icold - 1) * ng + (igold - 1)] += jr;
                if (p.itype == 1 && p.stat == 0) {
                    if (p.x > XA(p.ix + 1) || p.x < XA(p.ix) || p.x != p.x) {
                        if (ierr == 0) ierr = -99;
                    }
                }
                if (ierr < 0) ct[8]++;
                if (m->max_steps > 0 && istep >= m->max_steps && p.stat == 0 && ierr == 0) ierr = 200;
                if (ierr > 0) {
                    perr = ierr;
                    ct[7]++;
                    lerr = 1;
                    p.stat = 2;
                    break;
                }
            }
            if (istep > ct[3]) ct[3] = istep;
            ct[9] += p.nnewton;
            if (p.stat == 2) {
                ct[6]++;
                tot[4] += p.e;
            } else {
                ct[5]++;
                tot[2] -= p.e;
                tot[0] += p.e;
                int ig = or_binsrch(p.wl, m->wl, ng + 1, 0);   /* fluxtally.f:42 */
                lum[ig - 1] += p.e;
                dev[ig - 1] += p.e * p.e;
                num[ig - 1] += 1;
            }
            if (res->pstat) {
                res->pstat[ip] = (uint8_t)p.stat;
                res->px[ip] = p.x; res->pmu[ip] = p.mu; res->pwl[ip] = p.wl; res->pe[ip] = p.e;
                res->pig[ip] = p.ig; res->pix[ip] = p.ix; res->pnstep[ip] = istep; res->perr[ip] = perr;
            }
        }
#pragma omp critical
        {
            for (int i = 0; i < nc; ++i) res->tally[i] += tally[i];
            if (res->eamp)
                for (int i = 0; i < nc; ++i) res->eamp[i] += eamp[i];
            for (size_t k = 0; k < (size_t)ng * nc; ++k) res->jrad[k] += jrad[k];
            for (int g = 0; g < ng; ++g) {
                res->luminos[g] += lum[g];
                res->lumdev[g] += dev[g];
                res->lumnum[g] += (int)num[g];
            }
            for (int k = 0; k < 7; ++k) res->tot[k] += tot[k];
            for (int k = 0; k < 10; ++k) {
                if (k == 3) { if (ct[3] > res->ct[3]) res->ct[3] = ct[3]; }
                else res->ct[k] += ct[k];
            }
            if (lerr) any_err = 1;
        }
        free(tally); free(eamp); free(jrad); free(lum); free(dev); free(num);
        free(cache.specarr); free(cache.glumps); free(cache.llumps);
    }
    free(pstart);
    free(srcp);
    return any_err ? 1 : 0;
}

/* exported single Philox block, for the known-answer tests */
or_u32x4 or_philox_block(or_u32x4 c, uint32_t k0, uint32_t k1)
{
    return or_philox4x32_10(c, k0, k1);
}

/* exported stream draws, for the tests of the stream layout: n draws of the sequential
 * stream (step = 0) or of transport step `step` (step >= 1) of particle pid */
void or_stream_draws(uint32_t seed, uint32_t iter, uint32_t pid, uint32_t phase,
                     uint32_t step, int n, double *out)
{
    or_stream s;
    or_stream_init(&s, seed, iter, pid, phase);
    if (step > 0) or_stream_step(&s, step);
    for (int k = 0; k < n; ++k) out[k] = or_draw(&s);
}
