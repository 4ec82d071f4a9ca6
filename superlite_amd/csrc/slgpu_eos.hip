// LTE equation of state on gfx950 and its adjoint.
//
//   k_eos      GAS/gas_setup.f:101-160 (massfr2natomfr) + GAS/eos_update.f:23-45 +
//              ionsmod.f:160-356 (ions_solve_eos with saha_nelec): per cell the number
//              fractions, the Saha/secant solve for the electron fraction and the ground-level
//              populations of the context's ion list
//   k_eos_vjp  the adjoint of k_eos at its converged solution
//   k_eos_ndens the H I / H II number densities of the last solve (eos_update.f:47-58)
//
// One 64-lane wave runs one cell: lane z-1 takes element z's Saha recursion (and, for the
// populations, lane k the listed ion k), and every sum runs in the reference's serial order
// from LDS, so the results are those of the one-thread-per-cell order of operations; the
// partition functions are the host's, frozen at its first EOS call (ionsmod.f:208-215).
//
// Adjoint.  With u = nelec * ndens (the electron density; sahac2 = 2 sahac(T)/u), stage i of
// element z has the Saha weight log p_i = (i-1) log sahac2 - E_i/kT + log(q_i/q_1), E_i the
// sum of the ionisation energies below stage i.  The normalised fractions n_i = p_i / sum p
// give, with a_i = i-1 and <.> the n-weighted mean of the element,
//   d log n_i / d log u = -(a_i - <a>),
//   d log n_i / d T     = (1.5/T)(a_i - <a>) + (E_i - <E>)/(k T^2),
// and the element's mean charge ne_z = <a> has d ne_z/d log u = -Var(a),
// d ne_z/dT = (1.5/T) Var(a) + Cov(a, E)/(k T^2).  The solution satisfies
// F = nelec - sum_z natomfr_z ne_z(nelec*ndens, T) = 0 with dF/dnelec = 1 + sum_z fr_z Var_z/nelec,
// so the adjoint of nelec is one scalar division per cell (implicit function theorem).  The
// populations are those of the last Saha evaluation, as the forward stores them.
#include "slgpu_device.h"
#include "slgpu_internal.h"

namespace slgpu {

constexpr double PC_AMU = 1.66053886e-24;
constexpr double DBL_HUGE = 1.7976931348623157e308;
constexpr double DBL_TINY = 2.2250738585072014e-308;

#define ED_E(iz, ii) (d.eos_e[((iz) - 1) * 31 + (ii) - 1])
#define ED_Q(iz, ii) (d.eos_q[((iz) - 1) * 31 + (ii) - 1])
#define ED_G(iz, ii) (d.eos_g1[((iz) - 1) * 31 + (ii) - 1])

// ionsmod.f:311-354 for one element: the normalised stage fractions n[0..nion) at
// (kti, sahac2); false when the normalisation is invalid (the reference stops)
__device__ bool saha_element(const Dev &d, int iz, double kti, double sahac2, double *n)
{
    const int nion = d.eos_ni[iz - 1];
    double help = 0.0;
    for (int istart = 1; istart <= nion - 1; ++istart) {
        n[istart - 1] = 1.0;
        for (int ii = istart + 1; ii <= nion; ++ii) {
            help = sahac2 * exp(-kti * ED_E(iz, ii - 1)) * n[ii - 2] * ED_Q(iz, ii) / ED_Q(iz, ii - 1);
            n[ii - 1] = help;
        }
        if (help < DBL_HUGE) break;
    }
    double nsum = 0.0;
    for (int ii = 1; ii <= nion; ++ii) nsum += n[ii - 1];
    if (nsum != nsum || nsum > DBL_HUGE || nsum < DBL_TINY) return false;
    nsum = 1.0 / nsum;
    for (int ii = 1; ii <= nion; ++ii) n[ii - 1] = n[ii - 1] * nsum;
    return true;
}

__device__ bool saha_nelec(const Dev &d, const double *fr, double kti, double sahac2, double &out)
{
    double n[31];
    double ne_sum = 0.0;
    for (int iz = 1; iz <= 30; ++iz) {
        for (int k = 0; k < 31; ++k) n[k] = 0.0;
        if (!saha_element(d, iz, kti, sahac2, n)) return false;
        double ne = 0.0;
        for (int ii = 2; ii <= d.eos_ni[iz - 1]; ++ii) ne = ne + (ii - 1) * n[ii - 1];
        ne_sum = ne_sum + fr[iz - 1] * ne;
    }
    out = ne_sum;
    return true;
}

// flags: bit 0 a cell's solve stopped (nelec <= 0 or invalid Saha normalisation)
// One 64-lane wave per cell. Lane z-1 evaluates element z's Saha recursion (the 30
// recursions of one evaluation run in parallel); the wave then adds the 30 terms
// fr(z)*ne(z) from LDS in the reference's order with the serial loop's expression, so every lane holds the same sum, the
// secant steps run wave-uniform and the result is the serial one bit for bit.
__global__ __launch_bounds__(64) void k_eos(Dev d, const double *__restrict__ temp,
                                            const double *__restrict__ mass,
                                            const double *__restrict__ massfr, double acc,
                                            double *__restrict__ nelec, double *__restrict__ natom,
                                            double *__restrict__ natomfr, double *__restrict__ oc,
                                            double *__restrict__ ginv, int *__restrict__ niter)
{
    __shared__ double s_f[30], s_n[30];
    const int i = blockIdx.x, lane = threadIdx.x;
    if (i >= d.nx) return;
    double *fr = natomfr + (size_t)i * 30;
    if (lane == 0) {
        niter[i] = 0;
        d.eos_sahac2[i] = 0.0;
    }
    if (mass[i] <= 0.0) {
        for (int k = lane; k < d.nion; k += 64) {
            oc[(size_t)i * d.nion + k] = 0.0;
            ginv[(size_t)i * d.nion + k] = 0.0;
        }
        return;
    }
    // massfr2natomfr (gas_setup.f:125-157), no container fraction; every lane forms the same
    // serial sums, and the wave's fractions are kept in registers of every lane (lf)
    const double *mf = massfr + (size_t)i * 30;
    double help = 0.0;
    for (int l = 0; l < 30; ++l) help = help + mf[l];
    double na = 0.0;
    for (int l = 0; l < 30; ++l) na = na + (mf[l] / help) * mass[i] / (d.eos_amass[l] * PC_AMU);
    const int z = lane + 1;                                  // this lane's element (z <= 30)
    const double frz = z <= 30 ? ((mf[z - 1] / help) * mass[i] / (d.eos_amass[z - 1] * PC_AMU)) / na : 0.0;
    if (z <= 30) fr[z - 1] = frz;
    if (lane == 0) natom[i] = na;
    // ions_solve_eos (ionsmod.f:160-309)
    const double T = temp[i];
    const double ndens = na / d.vol[i];
    const double kti = 1.0 / (PC_KB * T);
    const double sahac = pow(PC_PI2 * PC_ME * PC_KB * T / (PC_H * PC_H), 1.5);
    double ne = nelec[i];
    double nel0 = 0.0, nel1 = 0.0, err0 = 1e10, err1 = -1e10, sahac2 = 0.0;
    int iprev = 0, iconv;
    for (iconv = 1; iconv <= 40; ++iconv) {
        sahac2 = 2.0 / (ne * ndens) * sahac;
        // saha_nelec: element z's term fr(z) * sum_ii (ii-1) n_ii on lane z-1
        bool ok = true;
        if (z <= 30) {
            double n[31];
            for (int k = 0; k < 31; ++k) n[k] = 0.0;
            ok = saha_element(d, z, kti, sahac2, n);
            double nez = 0.0;
            for (int ii = 2; ii <= d.eos_ni[z - 1]; ++ii) nez = nez + (ii - 1) * n[ii - 1];
            s_f[z - 1] = frz;
            s_n[z - 1] = nez;
        }
        const bool bad = __any(!ok);
        __syncthreads();
        if (bad) {
            if (lane == 0) atomicOr(d.flags, 1);
            return;
        }
        double ne_new = 0.0;
        for (int l = 0; l < 30; ++l) ne_new = ne_new + s_f[l] * s_n[l];
        __syncthreads();                                     // s_n is rewritten next evaluation
        const double err = ne / ne_new - 1.0;
        if (fabs(err) < acc) {
            if (iconv > 1) ne = ne_new;
            break;
        }
        if (err * err1 > err * err0) { nel1 = ne; err1 = err; }     // maxloc(err*nec%err)
        else { nel0 = ne; err0 = err; }
        if (iconv == 1) {
            double x = 1.5 * ne_new - 0.5 * ne;
            x = fmin(3.0 * ne_new, x);
            x = fmax(0.33 * ne_new, x);
            ne = x;
        } else {
            const int ih = (fabs(err1) > fabs(err0)) ? 2 : 1;
            const double h = fabs(err0 / err1);
            const double ynew = (ih != iprev && h < 1e5 && h > 1e-5) ? 0.0 : 0.001 * (ih == 1 ? err0 : err1);
            iprev = ih;
            const double dxdy = (nel1 - nel0) / (err1 - err0);
            ne = nel0 - dxdy * (err0 - ynew);
            if (ne <= 0.0) {
                if (lane == 0) atomicOr(d.flags, 1);
                return;
            }
        }
    }
    if (lane == 0) {
        nelec[i] = ne;
        niter[i] = iconv;
        d.eos_sahac2[i] = sahac2;
    }
    // eos_update.f:36-45: populations of the last Saha evaluation, one listed ion per lane
    for (int k = lane; k < d.nion; k += 64) {
        const int iz = d.ion_z[k], ii = d.ion_i[k];
        double n[31];
        for (int m = 0; m < 31; ++m) n[m] = 0.0;
        saha_element(d, iz, kti, sahac2, n);
        const double friz = ((mf[iz - 1] / help) * mass[i] / (d.eos_amass[iz - 1] * PC_AMU)) / na;
        ginv[(size_t)i * d.nion + k] = 1.0 / ED_G(iz, ii);
        oc[(size_t)i * d.nion + k] = na * friz * ED_G(iz, ii) * n[ii - 1] / (ED_Q(iz, ii) * d.vol[i]);
    }
}

// The adjoint of k_eos for one cell (see the header).  In: dL/dnatom, dL/dnatomfr (30),
// dL/dnelec, dL/doc (nion) of the outputs; out: dL/dT (the EOS's part), dL/dmass,
// dL/dmassfr (30).
__global__ __launch_bounds__(64) void k_eos_vjp(Dev d, const double *__restrict__ temp,
                                                const double *__restrict__ mass,
                                                const double *__restrict__ massfr,
                                                const double *__restrict__ nelec,
                                                const double *__restrict__ natom,
                                                const double *__restrict__ natomfr,
                                                const double *__restrict__ oc,
                                                const double *__restrict__ g_natom,
                                                const double *__restrict__ g_fr,
                                                const double *__restrict__ g_ne,
                                                const double *__restrict__ g_oc,
                                                double *__restrict__ out_T,
                                                double *__restrict__ out_mass,
                                                double *__restrict__ out_mf)
{
    // One 64-lane wave per cell: lane z-1 forms element z's stage moments, then lane k the
    // listed ion k's terms (dynamic LDS, 4 per ion), and the sums run in the serial order.
    extern __shared__ double s_ion[];
    __shared__ double s_ma[30], s_mE[30], s_var[30], s_cov[30], s_gfr[30];
    const int i = blockIdx.x, lane = threadIdx.x;
    if (i >= d.nx) return;
    if (lane == 0) {
        out_T[i] = 0.0;
        out_mass[i] = 0.0;
    }
    if (lane < 30) out_mf[(size_t)i * 30 + lane] = 0.0;
    if (mass[i] <= 0.0 || d.eos_sahac2[i] == 0.0) return;
    const double T = temp[i], ne = nelec[i], na = natom[i], V = d.vol[i];
    const double ndens = na / V;
    const double kti = 1.0 / (PC_KB * T);
    const double sahac2 = d.eos_sahac2[i];
    const double *fr = natomfr + (size_t)i * 30;
    const int nion = d.nion;
    // element z = lane + 1: the moments of its stage distribution
    if (lane < 30) {
        const int iz = lane + 1;
        double n[31];
        for (int m = 0; m < 31; ++m) n[m] = 0.0;
        saha_element(d, iz, kti, sahac2, n);
        const int ni = d.eos_ni[iz - 1];
        double ma = 0.0, ma2 = 0.0, mE = 0.0, maE = 0.0, Ei = 0.0;
        for (int ii = 1; ii <= ni; ++ii) {
            if (ii > 1) Ei += ED_E(iz, ii - 1);
            const double a = ii - 1;
            ma += a * n[ii - 1];
            ma2 += a * a * n[ii - 1];
            mE += Ei * n[ii - 1];
            maE += a * Ei * n[ii - 1];
        }
        s_ma[lane] = ma;
        s_mE[lane] = mE;
        s_var[lane] = ma2 - ma * ma;
        s_cov[lane] = maE - ma * mE;
    }
    __syncthreads();
    // listed ion k: its terms of dL/dnatom, dL/dfr(iz), the ln(u) and T adjoints
    for (int k = lane; k < nion; k += 64) {
        double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
        const double go = g_oc ? g_oc[(size_t)i * nion + k] : 0.0;
        if (go != 0.0) {
            const int iz = d.ion_z[k], ii = d.ion_i[k];
            double n[31];
            for (int m = 0; m < 31; ++m) n[m] = 0.0;
            saha_element(d, iz, kti, sahac2, n);
            double Ek = 0.0;
            for (int j = 1; j < ii; ++j) Ek += ED_E(iz, j);
            const double a = ii - 1, ma = s_ma[iz - 1], mE = s_mE[iz - 1];
            const double base = ED_G(iz, ii) * n[ii - 1] / (ED_Q(iz, ii) * V);   // oc / (natom fr)
            const double w = go * oc[(size_t)i * nion + k];
            t0 = go * fr[iz - 1] * base;
            t1 = go * na * base;
            t2 = -w * (a - ma);
            t3 = w * ((1.5 / T) * (a - ma) + (kti / T) * (Ek - mE));
        }
        s_ion[4 * k] = t0;
        s_ion[4 * k + 1] = t1;
        s_ion[4 * k + 2] = t2;
        s_ion[4 * k + 3] = t3;
    }
    __syncthreads();
    // the serial sums (the ion list ascends in iz), on every lane
    double gna = g_natom ? g_natom[i] : 0.0;
    double gT = 0.0, glu = 0.0, gnd = 0.0, sfv = 0.0, sfdT = 0.0;
    for (int l = 0; l < 30; ++l) {
        sfv += fr[l] * s_var[l];
        sfdT += fr[l] * ((1.5 / T) * s_var[l] + (kti / T) * s_cov[l]);
    }
    for (int k = 0; k < nion; ++k) {
        gna += s_ion[4 * k];
        glu += s_ion[4 * k + 2];
        gT += s_ion[4 * k + 3];
    }
    // u = nelec * ndens, then the implicit solution nelec(ndens, T, fr)
    double gne = g_ne ? g_ne[i] : 0.0;
    gne += glu / ne;
    gnd += glu / ndens;
    const double lam = gne / (1.0 + sfv / ne);
    gnd += -lam * sfv / ndens;
    gT += lam * sfdT;
    if (lane < 30) {
        double g = g_fr ? g_fr[(size_t)i * 30 + lane] : 0.0;
        for (int k = 0; k < nion; ++k)
            if (d.ion_z[k] == lane + 1) g += s_ion[4 * k + 1];
        s_gfr[lane] = g + lam * s_ma[lane];
    }
    __syncthreads();
    gna += gnd / V;
    // massfr2natomfr: fr = y/natom, natom = sum y, y_l = x_l mass/(A_l amu), x = mf/sum(mf)
    const double *mf = massfr + (size_t)i * 30;
    double H = 0.0;
    for (int l = 0; l < 30; ++l) H += mf[l];
    double sfg = 0.0;
    for (int l = 0; l < 30; ++l) sfg += s_gfr[l] * fr[l];
    const double gnat = gna - sfg / na;            // total adjoint of natom
    double sxg = 0.0, gm = 0.0, gxl = 0.0;
    for (int l = 0; l < 30; ++l) {
        const double gy = s_gfr[l] / na + gnat;
        const double c = mass[i] / (d.eos_amass[l] * PC_AMU);
        const double gx = gy * c;
        if (l == lane) gxl = gx;
        gm += gy * (mf[l] / H) / (d.eos_amass[l] * PC_AMU);
        sxg += gx * (mf[l] / H);
    }
    if (lane < 30) out_mf[(size_t)i * 30 + lane] = (gxl - sxg) / H;
    if (lane == 0) {
        out_mass[i] = gm;
        out_T[i] = gT;
    }
}

// eos_update.f:47-58 with in_nlte: nlte_ndens(i,1:2,1), the H I and H II number densities
// n_ii natom natomfr(1) / vol of the last solve (its last Saha evaluation, as the populations)
__global__ __launch_bounds__(64) void k_eos_ndens(Dev d, const double *__restrict__ temp,
                                                  const double *__restrict__ mass,
                                                  const double *__restrict__ natom,
                                                  const double *__restrict__ natomfr,
                                                  double *__restrict__ ndens)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.nx) return;
    double *o = ndens + 2 * (size_t)i;
    const double s2 = d.eos_sahac2[i];
    if (!(mass[i] > 0.0) || !(s2 > 0.0)) { o[0] = 0.0; o[1] = 0.0; return; }
    double n[31];
    for (int m = 0; m < 31; ++m) n[m] = 0.0;
    saha_element(d, 1, 1.0 / (PC_KB * temp[i]), s2, n);
    const double fr = natomfr[(size_t)i * 30];
    o[0] = n[0] * natom[i] * fr / d.vol[i];
    o[1] = n[1] * natom[i] * fr / d.vol[i];
}

void launch_eos_ndens(slgpu_ctx *c, const double *temp, const double *mass, const double *natom,
                      const double *natomfr, double *ndens)
{
    hipLaunchKernelGGL(k_eos_ndens, dim3((c->nx + 63) / 64), dim3(64), 0, c->stream, c->d, temp,
                       mass, natom, natomfr, ndens);
}

void launch_eos(slgpu_ctx *c, const double *temp, const double *mass, const double *massfr,
                double acc, double *nelec, double *natom, double *natomfr, double *oc,
                double *ginv, int *niter)
{
    hipLaunchKernelGGL(k_eos, dim3(c->nx), dim3(64), 0, c->stream, c->d, temp, mass,
                       massfr, acc, nelec, natom, natomfr, oc, ginv, niter);
}

void launch_eos_vjp(slgpu_ctx *c, const double *temp, const double *mass, const double *massfr,
                    const double *nelec, const double *natom, const double *natomfr,
                    const double *oc, const double *g_natom, const double *g_fr,
                    const double *g_ne, const double *g_oc, double *out_T, double *out_mass,
                    double *out_mf)
{
    hipLaunchKernelGGL(k_eos_vjp, dim3(c->nx), dim3(64), (size_t)4 * (c->nion > 0 ? c->nion : 1) * sizeof(double),
                       c->stream, c->d, temp,
                       mass, massfr, nelec, natom, natomfr, oc, g_natom, g_fr, g_ne, g_oc, out_T,
                       out_mass, out_mf);
}

}  // namespace slgpu
