// NLTE hydrogen opacity on gfx950 (GAS/physical_opacity_nlte.f:3-708 with nlte_nelem = 1).
//
//   k_nlte_rates   one block per massive cell: the H I rate coefficients (radiative with J
//                  at the group centre, :160-184; electron-impact excitation, :189-222),
//                  the photoionisation and recombination rates (:226-253) and the rate
//                  matrix (:255-266), column-major in the cell's slot of NlteDev::A
//   k_nlte_solve   one block per massive cell: LU with partial pivoting (LAPACK dgetf2's
//                  unblocked algorithm, rank-1 updates spread over the block) and the two
//                  substitutions of dgetrs; the normalised level populations
//   k_opacity_nlte one thread per (cell, group), 64 cells x 4 groups per block: the H lines
//                  of the group in the file's order (NLTE absorption with stimulated emission
//                  from the populations, :304-345), the other elements' lines in LTE
//                  (:349-401), bf (:413-452) and ff (:458-511), each added to cap, capemit and
//                  emiss as the reference adds them
//   k_nlte_finish  transpose to the (nc, ng) outputs; k_nlte_means the Planck mean of capemit
//                  (opacity_planckmean.f:22)
//   k_nlte_rad_est the radiation temperature and intensity estimates after the transport
//                  (nltemod.f:231-260)
//
// The reference inverts the rate matrix (dgetrf + dgetri) and multiplies; the populations
// here solve the same system with the same factorisation.  The arithmetic is FP64 in both
// precision forms, without fused multiply-add contraction except where the factorisation
// writes it explicitly (as the oracle does).
#include <float.h>
#include "slgpu_device.h"
#include "slgpu_internal.h"

#pragma clang fp contract(off)

namespace slgpu {

constexpr double KB_EV = PC_KB / PC_EV;

__global__ __launch_bounds__(256) void k_nlte_rates(Dev d, NlteDev q)
{
    const int i = blockIdx.x;
    const int n = q.n;
    if (!(d.mass[i] > 0.0)) {                             // uniform per block: no populations
        for (int k = threadIdx.x; k < n; k += blockDim.x) q.pop[(size_t)i * n + k] = 0.0;
        return;
    }
    double *rc = q.rc + (size_t)i * n * n;
    double *A = q.A + (size_t)i * n * n;
    double *rhs = q.pop + (size_t)i * n;
    double *gpi = q.gpi + (size_t)i * n;
    for (int k = threadIdx.x; k < n * n; k += blockDim.x) rc[k] = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) rhs[k] = k == 0 ? 1.0 : 0.0;
    __syncthreads();
    const double T = d.temp[i], tempinv = 1.0 / T;
    // radiative: rc(llw, lup) = Blu J, rc(lup, llw) = Aul + Bul J; the level pairs are
    // distinct (slgpu_set_nlte_data checks), so the assignments do not race
    for (int il = threadIdx.x; il < q.nlin; il += blockDim.x) {
        const int g = q.igs[il];
        if (g < 1) continue;
        const double wlm = 0.5 * (d.wl[g] + d.wl[g - 1]);
        const double jnu = (wlm * wlm / PC_C) * planck(wlm, T);
        const int lw = q.llw[il] - 1, lu = q.lup[il] - 1;
        rc[(size_t)lu * n + lw] = q.blu[il] * jnu;
        rc[(size_t)lw * n + lu] = q.aul[il] + q.bul[il] * jnu;
    }
    __syncthreads();
    // electron-impact excitation and de-excitation (distinct pairs as well)
    for (int il = threadIdx.x; il < q.ncoll; il += blockDim.x) {
        const double x = q.dele[il] * tempinv / KB_EV;
        const double lx = log(x);
        const double *c = q.cc + 4 * (size_t)il;
        double f = c[0] + c[1] * lx + c[2] * (lx * lx) + c[3] * (lx * lx * lx);
        f = exp(f);
        const double gamma = 1.58e-5 * (f / x) * exp(-1.0 * x) * pow(tempinv, 1.5) / pow(KB_EV, 1.5);
        const int lw = q.cl[il], lu = q.cu[il];
        const double gamma_d = q.glev[lw - 1] * gamma * exp(x) / q.glev[lu - 1];
        const double clu = gamma * d.nelec_in[i] * d.natom_in[i] / d.vol_in[i];
        const double cul = gamma_d * d.nelec_in[i] * d.natom_in[i] / d.vol_in[i];
        rc[(size_t)(lu - 1) * n + (lw - 1)] = rc[(size_t)(lu - 1) * n + (lw - 1)] + clu;
        rc[(size_t)(lw - 1) * n + (lu - 1)] = rc[(size_t)(lw - 1) * n + (lu - 1)] + cul;
    }
    // photoionisation and radiative recombination (each level once)
    for (int ll = threadIdx.x; ll < n; ll += blockDim.x) {
        const int l = q.levpi[ll];
        const double xpi = q.delepi[ll] * tempinv / KB_EV;
        const double lp = log(xpi);
        const double *cp = q.pic + 4 * (size_t)ll, *cr = q.rrc + 4 * (size_t)ll;
        double fpi = cp[0] + cp[1] * lp + cp[2] * (lp * lp) + cp[3] * (lp * lp * lp);
        fpi = exp(fpi);
        gpi[l - 1] = (fpi / xpi) * exp(-1.0 * xpi) * pow(tempinv, 1.5) / pow(KB_EV, 1.5);
        const double xrr = KB_EV * q.radtemp[i];
        const double lr = log(xrr);
        double frr = cr[0] + cr[1] * lr + cr[2] * (lr * lr) + cr[3] * (lr * lr * lr);
        frr = exp(frr);
        if (l > 1) rhs[l - 1] = -1.0 * frr * q.ndens[2 * (size_t)i + 1] / q.ndens[2 * (size_t)i];
    }
    __syncthreads();
    // the rate matrix: row 1 all ones; row l > 1: rc(ll, l) off the diagonal, and on it
    // -sum_k rc(l, k) - gamma_PI(l), the sum in index order as the Fortran sum takes it
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int r = e % n, col = e / n;
        double v;
        if (r == 0) v = 1.0;
        else if (r == col) {
            double sum = 0.0;
            for (int k = 0; k < n; ++k) sum += rc[(size_t)k * n + r];
            v = -1.0 * sum;
            v = v - gpi[r];
        } else v = rc[(size_t)r * n + col];
        A[e] = v;
    }
}

__global__ __launch_bounds__(256) void k_nlte_solve(Dev d, NlteDev q)
{
    const int i = blockIdx.x;
    if (!(d.mass[i] > 0.0)) return;
    const int n = q.n;
    double *a = q.A + (size_t)i * n * n;
    double *b = q.pop + (size_t)i * n;
    __shared__ double s_v[256];
    __shared__ int s_i[256];
    __shared__ int s_info;
    int *ipiv = q.ipiv + (size_t)i * n;
    if (threadIdx.x == 0) s_info = 0;
    const int t = threadIdx.x;
#define A_(r, c) a[(size_t)(c) * n + (r)]
    for (int j = 0; j < n; ++j) {
        // idamax over rows j..n-1: the first largest |a|
        {
            double best = -1.0;
            int bi = n;
            for (int r = j + t; r < n; r += blockDim.x) {
                const double v = fabs(A_(r, j));
                if (bi == n || v > best) { best = v; bi = r; }
            }
            s_v[t] = best;
            s_i[t] = bi;
            __syncthreads();
            for (int s = blockDim.x / 2; s > 0; s >>= 1) {
                if (t < s) {
                    const double v2 = s_v[t + s];
                    const int i2 = s_i[t + s];
                    if (i2 < n && (s_i[t] == n || v2 > s_v[t] || (v2 == s_v[t] && i2 < s_i[t]))) {
                        s_v[t] = v2;
                        s_i[t] = i2;
                    }
                }
                __syncthreads();
            }
        }
        const int jp = s_i[0];
        const bool nz = A_(jp, j) != 0.0;
        __syncthreads();
        if (t == 0) ipiv[j] = jp;
        if (nz) {
            if (jp != j)
                for (int k = t; k < n; k += blockDim.x) {
                    const double tmp = A_(j, k);
                    A_(j, k) = A_(jp, k);
                    A_(jp, k) = tmp;
                }
            __syncthreads();
            const double piv = A_(j, j);
            if (fabs(piv) >= DBL_MIN) {
                const double rcp = 1.0 / piv;
                for (int r = j + 1 + t; r < n; r += blockDim.x) A_(r, j) = A_(r, j) * rcp;
            } else {
                for (int r = j + 1 + t; r < n; r += blockDim.x) A_(r, j) = A_(r, j) / piv;
            }
        } else if (t == 0 && s_info == 0) {
            s_info = j + 1;
        }
        __syncthreads();
        const int m = n - j - 1;
        for (int e = t; e < m * m; e += blockDim.x) {
            const int r = j + 1 + e % m, k = j + 1 + e / m;
            const double u = A_(j, k);
            if (u != 0.0) A_(r, k) = fma(-A_(r, j), u, A_(r, k));
        }
        __syncthreads();
    }
    if (t == 0) {
        if (s_info) {
            atomicOr(q.flag, 1);
        } else {
            for (int k = 0; k < n; ++k)
                if (ipiv[k] != k) { const double tmp = b[k]; b[k] = b[ipiv[k]]; b[ipiv[k]] = tmp; }
            for (int k = 0; k < n; ++k)
                if (b[k] != 0.0)
                    for (int r = k + 1; r < n; ++r) b[r] = fma(-b[k], A_(r, k), b[r]);
            for (int k = n - 1; k >= 0; --k)
                if (b[k] != 0.0) {
                    b[k] = b[k] / A_(k, k);
                    for (int r = 0; r < k; ++r) b[r] = fma(-b[k], A_(r, k), b[r]);
                }
        }
    }
#undef A_
}

__global__ __launch_bounds__(256) void k_opacity_nlte(Dev d, NlteDev q, int flags, double ffc1,
                                                      const double *__restrict__ rgg_tab)
{
    const int nc = d.nx, ng = d.ng, n = q.n;
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 64 + lane;
    const int g = __builtin_amdgcn_readfirstlane(blockIdx.y * 4 + (threadIdx.x >> 6)) + 1;
    if (g > ng) return;
    const bool act = i < nc;
    const int ic = act ? i : nc - 1;
    const bool massive = act && d.mass[ic] > 0.0;
    const double T = d.temp[ic], hckt = d.hckt[ic];
    const double wlo = d.wl[g - 1], whi = d.wl[g];
    const double fconst = PC_PI * (PC_E * PC_E) / (PC_ME * PC_C);
    double cap = 0.0, cape = 0.0, em = 0.0;
    if (!(flags & SLGPU_NOBB)) {
        const double wl = 0.5 * (wlo + whi);
        const double dwl = whi - wlo;
        const double phi = wl * wl / (dwl * PC_C);
        const double bnu = (wl * wl / PC_C) * planck(wl, T);
        if (massive) {
            const double nd = q.ndens[2 * (size_t)ic];
            const double *pop = q.pop + (size_t)ic * n;
            for (int h = q.hstart[g]; h < q.hstart[g + 1]; ++h) {
                const int il = q.hidx[h];
                const double npl = pop[q.llw[il] - 1] * nd;
                const double npu = pop[q.lup[il] - 1] * nd;
                const double caphelp = npl * fconst * q.f[il] * phi * (1.0 - (npu * q.gl[il] / (npl * q.gu[il])));
                cap += caphelp;
                const double emithelp = npu * q.aul[il] * phi * PC_H * PC_C / (4.0 * PC_PI * (q.wl0[il] * PC_ANG));
                cape += emithelp / bnu;
                em += emithelp;
            }
        }
        const double wlinv = 1.0 / wl;
        const double expfac = 1.0 - exp(-hckt * wlinv);
        const double *ocg = d.ocg + ic;
        for (int l = d.gstart[g]; l < d.gstart[g + 1]; ++l) {
            const LineRec r = d.lines[l];                   // uniform: scalar loads
            if (r.ion == q.kH) continue;                    // H I is the NLTE species
            const double o = ocg[(size_t)r.ion * nc];
            if (!(o > 0.0) || !massive) continue;
            const double caphelp = phi * (double)r.gxs * o * exp(-(double)r.chilw * hckt) * expfac;
            cap += caphelp;
            cape += caphelp;
            em += caphelp * bnu;
        }
    }
    const double bnul = (wlo * wlo / PC_C) * planck(wlo, T);     // bf and ff: left edge
    if (!(flags & SLGPU_NOBF) && massive) {
        const double *xs = d.bfxs + (size_t)(g - 1) * d.nion;
        for (int k = 0; k < d.nion; ++k) {
            const double x = xs[k];
            if (x != 0.0) {
                const double v = x * d.ocT[(size_t)k * nc + ic];
                cap += v;
                cape += v;
                em += v * bnul;
            }
        }
    }
    if (!(flags & SLGPU_NOFF) && massive) {
        const double wl = wlo;
        const double u = hckt * (1.0 / wl);
        const double ru = 10.0 * (log10(u) + 4.0) + 1.0;
        int iu = (int)floor(ru);
        iu = iu < 1 ? 1 : (iu > FF_NU - 1 ? FF_NU - 1 : iu);
        const double du = ru - iu;
        const double help = ffc1 * sqrt(hckt) * (1.0 - exp(-u)) * (wl * wl * wl) * d.hlparr[ic];
        const double v = help * ff_gaunt_sum(d, ic, iu, du, rgg_tab);
        cap += v;
        cape += v;
        em += v * bnul;
    }
    if (act) {
        const size_t o = (size_t)(g - 1) * nc + i;
        d.capT[o] = massive ? cap : 0.0;
        q.capemitT[o] = massive ? cape : 0.0;
        q.emissT[o] = massive ? em : 0.0;
    }
}

// (ng, nc) cell-fastest -> (nc, ng) group-fastest for cap, capemit and emiss
__global__ __launch_bounds__(256) void k_nlte_finish(Dev d, NlteDev q, int want64)
{
    __shared__ double t0[32][33], t1[32][33], t2[32][33];
    const int nc = d.nx, ng = d.ng;
    const int g0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int r = ty; r < 32; r += 8) {
        const int g = g0 + r, i = i0 + tx;
        const bool in = g < ng && i < nc;
        const size_t o = (size_t)g * nc + i;
        t0[r][tx] = in ? d.capT[o] : 0.0;
        t1[r][tx] = in ? q.capemitT[o] : 0.0;
        t2[r][tx] = in ? q.emissT[o] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, g = g0 + tx;
        if (i < nc && g < ng) {
            const size_t o = (size_t)i * ng + g;
            if (want64) d.cap64[o] = t0[tx][r];
            d.cap32[o] = (float)t0[tx][r];
            q.capemit32[o] = (float)t1[tx][r];
            d.emiss[o] = (float)t2[tx][r];
        }
    }
}

// Planck mean of the FP32 capemit with the specarr of k_cell_tables, in its reduction order
__global__ __launch_bounds__(256) void k_nlte_means(Dev d, NlteDev q)
{
    __shared__ double sh[4];
    const int i = blockIdx.x, ng = d.ng;
    double pe = 0.0;
    for (int g = threadIdx.x + 1; g <= ng; g += 256)
        pe += (double)q.capemit32[(size_t)i * ng + g - 1] * d.specarr[(size_t)i * ng + g - 1];
    pe = block_sum256(pe, sh);
    if (threadIdx.x == 0) q.capemitgrey[i] = pe;
}

// nlte_rad_est (nltemod.f:231-260) after the transport: the radiation temperature estimate
// (tally/(a V))^(1/4) from the census energy tally, damped by one half, and the group-centre
// intensity wlm^2 B(wlm, radtemp)/c at the new radiation temperature; one block per cell
__global__ __launch_bounds__(256) void k_nlte_rad_est(Dev d, double *radtemp, double *jrad)
{
    __shared__ double s_tr;
    const int i = blockIdx.x, ng = d.ng;
    if (threadIdx.x == 0) {
        const double trad = pow(d.tally[i] / (PC_ACOEF * d.vol[i]), 0.25);
        const double r = radtemp[i];
        const double rn = r + 0.5 * (trad - r);
        radtemp[i] = rn;
        s_tr = rn;
    }
    __syncthreads();
    if (!jrad) return;
    const double tr = s_tr;
    for (int g = threadIdx.x + 1; g <= ng; g += 256) {
        const double wlm = 0.5 * (d.wl[g] + d.wl[g - 1]);
        jrad[(size_t)i * ng + g - 1] = wlm * wlm * planck(wlm, tr) / PC_C;
    }
}

void launch_nlte_rad_est(slgpu_ctx *c, double *radtemp, double *jrad)
{
    hipLaunchKernelGGL(k_nlte_rad_est, dim3(c->nx), dim3(256), 0, c->stream, c->d, radtemp, jrad);
}

void launch_opacity_nlte(slgpu_ctx *c, int flags, double ffc1, const double *rgg)
{
    const NlteDev &q = c->nl;
    if (!(flags & SLGPU_NOBB)) {
        hipLaunchKernelGGL(k_nlte_rates, dim3(c->nx), dim3(256), 0, c->stream, c->d, q);
        hipLaunchKernelGGL(k_nlte_solve, dim3(c->nx), dim3(256), 0, c->stream, c->d, q);
    }
    dim3 grid((c->nx + 63) / 64, (c->ng + 3) / 4);
    hipLaunchKernelGGL(k_opacity_nlte, grid, dim3(256), 0, c->stream, c->d, q, flags, ffc1, rgg);
    dim3 tg((c->ng + 31) / 32, (c->nx + 31) / 32);
    hipLaunchKernelGGL(k_nlte_finish, tg, dim3(256), 0, c->stream, c->d, q, 1);
}

void launch_nlte_means(slgpu_ctx *c)
{
    hipLaunchKernelGGL(k_nlte_means, dim3(c->nx), dim3(256), 0, c->stream, c->d, c->nl);
}

}  // namespace slgpu
