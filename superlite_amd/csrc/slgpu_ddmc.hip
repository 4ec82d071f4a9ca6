// DDMC preparation on gfx950: lumped leakage opacities (GRID/leakage_opacity11.f90:30-171),
// binned emission probabilities (GRID/emission_probability.f90:39-62), and the per-cell
// sampling tables that replace the reference's per-event O(ng) linear scans
// (GRID/emitgroup.f:23-51, TRANSPORT1/diffusion11.f90:289-313, 375-398, 447-460,
// 489-511, 571-584; SOURCE/interior_source.f90:96-103) by cumulative rows plus exact
// bucket maps (slgpu_sample.h), so a sampled group usually costs one table load.  One
// 1024-thread block per cell; groups are swept in chunks of 1024 with block-wide (wave64
// shuffle) scans carrying the running prefix.
#include "slgpu_device.h"
#include "slgpu_internal.h"
#include "slgpu_sample.h"

namespace slgpu {

constexpr int NT = 1024;
constexpr int NW = NT / 64;

// inclusive block scan (fixed order) returning the block total through `total`
__device__ __forceinline__ double block_scan(double v, double *sh, double &total)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double x = wave_inclusive_scan(v);
    if (lane == 63) sh[w] = x;
    __syncthreads();
    if (w == 0) {
        double t = lane < NW ? sh[lane] : 0.0;
        t = wave_inclusive_scan(t);
        if (lane < NW) sh[lane] = t;
    }
    __syncthreads();
    double off = w > 0 ? sh[w - 1] : 0.0;
    total = sh[NW - 1];
    __syncthreads();
    return x + off;
}

__device__ __forceinline__ int block_scan_int(int v, int *sh, int &total)
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int x = v;
    #pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(x, off, 64);
        if (lane >= off) x += t;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    if (w == 0) {
        int t = lane < NW ? sh[lane] : 0;
        #pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            int u = __shfl_up(t, off, 64);
            if (lane >= off) t += u;
        }
        if (lane < NW) sh[lane] = t;
    }
    __syncthreads();
    int off = w > 0 ? sh[w - 1] : 0;
    total = sh[NW - 1];
    __syncthreads();
    return x + off;
}

// in-place running sum of row[0..n) in index order (loads batched, adds sequential)
__device__ void running_sum(double *row, int n)
{
    double s = 0.0;
    int j = 0;
    for (; j + 8 <= n; j += 8) {
        double v[8];
        #pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = row[j + k];
        #pragma unroll
        for (int k = 0; k < 8; ++k) { s = s + v[k]; row[j + k] = s; }
    }
    for (; j < n; ++j) { s = s + row[j]; row[j] = s; }
}

// sum of row[0..n) in index order
__device__ double sequential_sum(const double *row, int n)
{
    double s = 0.0;
    int j = 0;
    for (; j + 8 <= n; j += 8) {
        double v[8];
        #pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = row[j + k];
        #pragma unroll
        for (int k = 0; k < 8; ++k) s = s + v[k];
    }
    for (; j < n; ++j) s = s + row[j];
    return s;
}

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    double t;
    block_scan(v, sh, t);
    return t;
}

#define CAPC(ig, ic) ((double)d.cap32[(size_t)((ic) - 1) * ng + ((ig) - 1)])
// emission opacity: grd_capemit after an NLTE build, grd_cap after an LTE one
#define CAPE(ig, ic) ((double)d.capem32[(size_t)((ic) - 1) * ng + ((ig) - 1)])
#define XA(i) (d.xarr[(i) - 1])
#define DX(i) (d.xarr[(i)] - d.xarr[(i) - 1])
#define DX3(i) (d.xarr[(i)] * d.xarr[(i)] * d.xarr[(i)] - d.xarr[(i) - 1] * d.xarr[(i) - 1] * d.xarr[(i) - 1])
#define SIGC(ic) (d.sig[(ic) - 1])

// resolved leakage opacity of group g through the left (side<0) or right face of cell ix
// (diffusion11.f90:204-239, 383-395, 455-457, 496-508)
__device__ __forceinline__ double resleak(const Dev &d, int ix, int g, int side)
{
    const int ng = d.ng, ic = ix;
    if (side < 0) {
        const int l = ix - 1;
        if ((CAPC(g, l) + SIGC(l)) * DX(ix - 1) < d.tauddmc) {
            double mfp = (CAPC(g, ic) + SIGC(ic)) * DX(ix);
            double ppl = 4.0 / (3.0 * mfp + 6.0 * PC_DEXT);
            return 1.5 * ppl * (XA(ix) * XA(ix)) / DX3(ix);
        }
        double mfp = (SIGC(ic) + CAPC(g, ic)) * DX(ix) + (SIGC(l) + CAPC(g, l)) * DX(ix - 1);
        return 2.0 * (XA(ix) * XA(ix)) / (mfp * DX3(ix));
    }
    if (ix == d.nx) {
        double mfp = (CAPC(g, ic) + SIGC(ic)) * DX(ix);
        double ppr = 4.0 / (3.0 * mfp + 6.0 * PC_DEXT);
        return 1.5 * ppr * (XA(ix + 1) * XA(ix + 1)) / DX3(ix);
    }
    const int l = ix + 1;
    if ((CAPC(g, l) + SIGC(l)) * DX(ix + 1) < d.tauddmc) {
        double mfp = (CAPC(g, ic) + SIGC(ic)) * DX(ix);
        double ppr = 4.0 / (3.0 * mfp + 6.0 * PC_DEXT);
        return 1.5 * ppr * (XA(ix + 1) * XA(ix + 1)) / DX3(ix);
    }
    double mfp = (SIGC(ic) + CAPC(g, ic)) * DX(ix) + (SIGC(l) + CAPC(g, l)) * DX(ix + 1);
    return 2.0 * (XA(ix + 1) * XA(ix + 1)) / (mfp * DX3(ix));
}

// first index of the block's maximum of (v, k) pairs, k = 0 meaning none (the strict '>' of
// an ascending loop): the result is valid on thread 0
__device__ __forceinline__ void block_argmax_first(double &v, int &k, double *s_v, int *s_k)
{
    for (int off = 32; off > 0; off >>= 1) {
        double v2 = __shfl_xor(v, off, 64);
        int k2 = __shfl_xor(k, off, 64);
        if (v2 > v || (v2 == v && k2 != 0 && (k == 0 || k2 < k))) { v = v2; k = k2; }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { s_v[w] = v; s_k[w] = k; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double bv = 0.0;
        int bk = 0;
        for (int q = 0; q < NW; ++q)
            if (s_v[q] > bv || (s_v[q] == bv && s_k[q] != 0 && (bk == 0 || s_k[q] < bk))) {
                bv = s_v[q]; bk = s_k[q];
            }
        v = bv;
        k = bk;
    }
    __syncthreads();
}

__global__ __launch_bounds__(NT) void k_leakage(Dev d)
{
    __shared__ double shd[NW];
    __shared__ int shi[NW];
    __shared__ double s_emax[NW];
    __shared__ int s_imax[NW];
    const int ng = d.ng, nx = d.nx;
    const int i = blockIdx.x + 1, l = i;
    const int icnb1 = i - 1 < 1 ? 1 : i - 1, icnb2 = i + 1 > nx ? nx : i + 1;
    const double dist = DX(i);
    const double tinv = d.tempinv[l - 1];
    const double sg = SIGC(l);
    const double *spec = d.specarr + (size_t)(l - 1) * ng;
    const double divv = d.divv[l - 1];
    // pass 1: speclump and the lump size (leakage_opacity11.f90:52-59)
    double sl = 0.0;
    int nl = 0;
    for (int g = threadIdx.x + 1; g <= ng; g += NT) {
        const double c = CAPC(g, l);
        if (c * dist >= d.taulump && (sg + c) * dist >= d.tauddmc) { sl += spec[g - 1]; nl++; }
    }
    sl = block_sum(sl, shd);
    int nlump;
    block_scan_int(nl, shi, nlump);
    const double speclump = sl > 0.0 ? 1.0 / sl : 0.0;
    // pass 2: caplump, capemitlump (NLTE), igemitmax, doplump, leakage opacities (:61-168)
    double caplump = 0.0, capemitlump = 0.0, doplump = 0.0, o1 = 0.0, o2 = 0.0, emax = 0.0, emax2 = 0.0;
    int imax = 0, imax2 = 0;
    const double dvol = 3.0 * d.vol[l - 1] / PC_PI4;
    for (int g = threadIdx.x + 1; g <= ng; g += NT) {
        const double c = CAPC(g, l);
        const bool lump = !(c * dist < d.taulump) && !((sg + c) * dist < d.tauddmc);
        if (lump) {
            const double h = spec[g - 1] * c;
            caplump += h;
            if (h > emax) { emax = h; imax = g; }
            if (d.nlte) {
                const double he = spec[g - 1] * CAPE(g, l);
                capemitlump += he;
                if (he > emax2) { emax2 = he; imax2 = g; }
            }
            const double specval = spec[g - 1];
            bool lh = (i == 1) ? true : ((CAPC(g, icnb1) + SIGC(icnb1)) * DX(i - 1) < d.tauddmc);
            if (lh) {
                double help = (c + sg) * DX(i);
                double ppl = 4.0 / (3.0 * help + 6.0 * PC_DEXT);
                o1 += (specval * speclump) * 1.5 * ppl * (XA(i) * XA(i)) / dvol;
            } else {
                double help = (sg + c) * DX(i) + (SIGC(icnb1) + CAPC(g, icnb1)) * DX(i - 1);
                o1 += (specval * speclump) * 2.0 * (XA(i) * XA(i)) / (help * 3.0 * d.vol[l - 1] / PC_PI4);
            }
            lh = (i == nx) ? true : ((CAPC(g, icnb2) + SIGC(icnb2)) * DX(i + 1) < d.tauddmc);
            if (lh) {
                double help = (c + sg) * DX(i);
                double ppr = 4.0 / (3.0 * help + 6.0 * PC_DEXT);
                o2 += (specval * speclump) * 1.5 * ppr * (XA(i + 1) * XA(i + 1)) / dvol;
            } else {
                double help = (sg + c) * DX(i) + (SIGC(icnb2) + CAPC(g, icnb2)) * DX(i + 1);
                o2 += (specval * speclump) * 2.0 * (XA(i + 1) * XA(i + 1)) / (help * 3.0 * d.vol[l - 1] / PC_PI4);
            }
            if (divv >= 0) {
                if (g <= ng - 1 && !(CAPC(g + 1, l) * dist >= d.taulump))
                    doplump += dopspeccalc(tinv, d.wlinv, g) * divv / (3 * PC_C);
            } else {
                if (g >= 2 && !(CAPC(g - 1, l) * dist >= d.taulump))
                    doplump -= dopspeccalc(tinv, d.wlinv, g - 1) * divv / (3 * PC_C);
            }
        }
    }
    caplump = block_sum(caplump, shd);
    doplump = block_sum(doplump, shd);
    o1 = block_sum(o1, shd);
    o2 = block_sum(o2, shd);
    // igemitmax: first index of the maximum (strict '>' in the reference's ascending loop);
    // in NLTE the capemit loop continues from the cap loop's maximum, so its first maximum
    // wins only where it is strictly larger
    block_argmax_first(emax, imax, s_emax, s_imax);
    if (d.nlte) {
        capemitlump = block_sum(capemitlump, shd);
        block_argmax_first(emax2, imax2, s_emax, s_imax);
        if (emax2 > emax) imax = imax2;
    }
    double *op = d.opaclump + (size_t)(l - 1) * 11;
    const double capgrey = d.capgrey[l - 1];
    // capgreyinv (LTE) or capemitgreyinv (NLTE), diffusion11.f90:98-99
    double cgi = 1.0 / d.capemgrey[l - 1];
    if (!(cgi >= 0.0)) cgi = 0.0;
    const int nl_eff = speclump == 0.0 ? 0 : nlump;
    double c_emitlump, c_caplump, c_doplump;
    if (nl_eff == ng) {
        c_emitlump = 1.0; c_caplump = capgrey; c_doplump = 0.0;
    } else {
        c_caplump = caplump * speclump;
        double el = (d.nlte ? capemitlump : caplump) * cgi;     // diffusion11.f90:136-141
        c_emitlump = el < 1.0 ? el : 1.0;
        c_doplump = doplump * speclump;
    }
    if (threadIdx.x == 0) {
        op[0] = o1; op[1] = o2; op[2] = 0.0; op[3] = 0.0; op[4] = 0.0; op[5] = 0.0;
        op[6] = speclump; op[7] = caplump; op[8] = d.nlte ? capemitlump : 0.0;
        op[9] = (double)imax; op[10] = doplump;
        CellDDMC cc;
        cc.emgreyinv = cgi; cc.speclump = speclump; cc.emitlump = c_emitlump;
        cc.caplump = c_caplump; cc.doplump = c_doplump; cc.nlump = nl_eff; cc.pad = 0;
        d.cell[l - 1] = cc;
    }
    if (nl_eff == 0) return;

    // pass 3: lump order and cumulative sampling weights
    int *gorder = d.gorder + (size_t)(l - 1) * ng;
    int baseL = 0, baseU = nl_eff;
    for (int g0 = 1; g0 <= ng; g0 += NT) {
        const int g = g0 + threadIdx.x;
        int lump = 0;
        if (g <= ng) {
            const double c = CAPC(g, l);
            lump = (c * dist >= d.taulump && (sg + c) * dist >= d.tauddmc) ? 1 : 0;
        }
        int tl, tu;
        int pl = block_scan_int(lump, shi, tl);
        int pu = block_scan_int(g <= ng ? 1 - lump : 0, shi, tu);
        if (g <= ng) gorder[lump ? baseL + pl - 1 : baseU + pu - 1] = g;
        baseL += tl;
        baseU += tu;
    }
    __syncthreads();
    // a face or Doppler channel with zero total opacity is never chosen (its event
    // probability is 0); its row is kept finite (all zero) so every row is monotone
    const double hL = o1 > 0.0 ? 1.0 / o1 : 0.0, hR = o2 > 0.0 ? 1.0 / o2 : 0.0;
    const double hD = c_doplump > 0.0 ? 1.0 / c_doplump : 0.0;
    const double d2s = 1.0 / (1.0 - c_emitlump);
    double *cL = d.cumL + (size_t)(l - 1) * ng, *cR = d.cumR + (size_t)(l - 1) * ng;
    double *cD = d.cumD + (size_t)(l - 1) * ng, *cS = d.cumS + (size_t)(l - 1) * ng;
    // terms in parallel, then each row's running sum in index order by one lane: the
    // reference accumulates sequentially, and a sequential sum of nonnegative terms is
    // monotone (a parallel scan is not, by rounding), which the samplers' searches need
    for (int p = threadIdx.x; p < nl_eff; p += NT) {
        // the Doppler scan runs ascending for redshift, descending for blueshift
        const int pd = divv >= 0 ? p : nl_eff - 1 - p;
        const int g = gorder[p], gd = gorder[pd];
        const double sp = spec[g - 1];
        double tD = 0.0;
        cL[p] = i > 1 ? sp * resleak(d, i, g, -1) * speclump * hL : 0.0;
        cR[p] = sp * resleak(d, i, g, +1) * speclump * hR;
        if (divv >= 0) {
            if (gd != ng && !(CAPC(gd + 1, l) * dist >= d.taulump))
                tD = dopspeccalc(tinv, d.wlinv, gd) * divv / (3 * PC_C) * speclump * hD;
        } else {
            if (gd != 1 && !(CAPC(gd - 1, l) * dist >= d.taulump))
                tD = -(dopspeccalc(tinv, d.wlinv, gd - 1) * divv / (3 * PC_C) * speclump * hD);
        }
        cD[p] = tD;
    }
    for (int p = nl_eff + threadIdx.x; p < ng; p += NT) {
        const int g = gorder[p];
        cS[p] = spec[g - 1] * CAPE(g, l) * cgi * d2s;     // diffusion11.f90:574-581
    }
    __syncthreads();
    if (threadIdx.x == 0) running_sum(cL, nl_eff);
    else if (threadIdx.x == 64) running_sum(cR, nl_eff);
    else if (threadIdx.x == 128) running_sum(cD, nl_eff);
    else if (threadIdx.x == 192) running_sum(cS + nl_eff, ng - nl_eff);
}

// emission_probability.f90:39-62 (grd_capemit, grd_capemitgrey in NLTE) and the in-bin
// cumulative sums emitgroup.f:38-50 walks
__global__ __launch_bounds__(256) void k_emitprob(Dev d)
{
    __shared__ double binsum[1024];
    __shared__ double pref[1025];
    const int ng = d.ng, nep = d.nep, nepg = d.nepg;
    const int l = blockIdx.x + 1;
    const double *spec = d.specarr + (size_t)(l - 1) * ng;
    for (int b = threadIdx.x; b < nep; b += 256) {
        const int g0 = b * nepg + 1;
        const int g1 = (b + 1) * nepg < ng ? (b + 1) * nepg : ng;
        double s = 0.0;
        for (int g = g0; g <= g1; ++g) s += spec[g - 1] * CAPE(g, l);
        binsum[b] = s;
    }
    __syncthreads();
    double *ep = d.emitprob + (size_t)(l - 1) * nep;
    if (threadIdx.x == 0) {
        double help = 0.0;
        pref[0] = 0.0;
        for (int iep = 1; iep <= nep - 1; ++iep) {
            help = help + binsum[iep - 1];
            ep[iep - 1] = help;
            pref[iep] = help;
        }
        ep[nep - 1] = d.capemgrey[l - 1];
    }
    __syncthreads();
    double *cum = d.emitcum + (size_t)(l - 1) * ng;
    for (int b = threadIdx.x; b < nep; b += 256) {
        const int g0 = b * nepg + 1;
        const int g1 = (b + 1) * nepg < ng ? (b + 1) * nepg : ng;
        double s = pref[b];
        for (int g = g0; g <= g1; ++g) {
            s = s + spec[g - 1] * CAPE(g, l);
            cum[g - 1] = s;
        }
    }
}

// normalised Planck at T_rad and its cumulative (interior_source.f90:69-75, 96-103):
// terms in parallel; the normalising sum and the running sum in group order, as the
// reference's loops do
__global__ __launch_bounds__(NT) void k_srccum(Dev d)
{
    __shared__ double s_norm;
    const int ng = d.ng;
    const int l = blockIdx.x;
    const double tr = 1.0 / d.radtemp[l];
    double *cum = d.srccum + (size_t)l * ng;
    for (int g = threadIdx.x + 1; g <= ng; g += NT) cum[g - 1] = specintv_g(tr, d.wlinv, g);
    __syncthreads();
    if (threadIdx.x == 0) s_norm = sequential_sum(cum, ng);
    __syncthreads();
    const double s = s_norm;
    for (int g = threadIdx.x + 1; g <= ng; g += NT) cum[g - 1] = cum[g - 1] / s;
    __syncthreads();
    if (threadIdx.x == 0) running_sum(cum, ng);
}

// exact bucket maps of the five group samplers; one thread per (bucket, cell, table)
__global__ __launch_bounds__(256) void k_bucket_tables(Dev d)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    const int ic = blockIdx.y + 1, t = blockIdx.z;
    if (b >= BT_K) return;
    uint32_t *row = d.btab + ((size_t)t * d.nx + (ic - 1)) * BT_K;
    const double rlo = bt_lo(b), rhi = bt_hi(b);
    uint32_t e = 0;
    if (t == TAB_E) {
        const int glo = emitgroup_direct(d, rlo, ic), ghi = emitgroup_direct(d, rhi, ic);
        if (glo == ghi) e = BT_PURE | dest_flags(d, t, ic, glo) | (uint32_t)glo;
        else e = ((uint32_t)ghi << 15) | (uint32_t)glo;
    } else {
        const LumpRow L = lump_row(d, t, ic, d.cell[ic - 1].nlump);
        if (L.cnt > 0) {                    // rows of cells without lumping are never sampled
            const int jlo = first_above(L.cum, 0, L.cnt, rlo);
            const int jhi = first_above(L.cum, 0, L.cnt, rhi);
            const int plo = jlo < L.cnt ? jlo : L.cnt - 1, phi = jhi < L.cnt ? jhi : L.cnt - 1;
            if (plo == phi) {
                const int g = lump_group(d, L, ic, jlo);
                e = BT_PURE | dest_flags(d, t, ic, g) | (uint32_t)g;
            } else {
                e = ((uint32_t)jhi << 15) | (uint32_t)jlo;
            }
        }
    }
    row[b] = e;
}

void launch_ddmc_tables(slgpu_ctx *c)
{
    hipLaunchKernelGGL(k_leakage, dim3(c->nx), dim3(NT), 0, c->stream, c->d);
    hipLaunchKernelGGL(k_emitprob, dim3(c->nx), dim3(256), 0, c->stream, c->d);
    hipLaunchKernelGGL(k_bucket_tables, dim3(BT_K / 256, c->nx, NTAB), dim3(256), 0, c->stream,
                       c->d);
    hipLaunchKernelGGL(k_srccum, dim3(c->nx), dim3(NT), 0, c->stream, c->d);
}

}  // namespace slgpu
