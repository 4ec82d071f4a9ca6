// LTE group-opacity build on gfx950 (GAS/physical_opacity.f:3-357, grp_ngs == 1) and the
// Planck / Rosseland means (GAS/opacity_planckmean.f:18-24, GAS/opacity_rossmean.f:18-21).
//
// Work decomposition.  One thread per (cell, group).  A wave holds 64 consecutive cells
// of ONE group, so the group's slice of the lambda-sorted line list is wave-uniform: the
// line records stream through the scalar unit (s_load) while the 64 lanes evaluate the
// per-cell Boltzmann factor exp(-chi_low*hc/kT) in parallel.  Each (cell, group) sum is a
// segment of the segmented reduction over the sorted lines; the segment boundaries are
// the precomputed per-group line offsets, so every sum runs in the reference's line order
// and no cross-lane combine is needed.  bf and ff are accumulated into the same register
// afterwards, in the reference's order (bb, then bf by ion, then ff by element).
#include <type_traits>
#include "slgpu_device.h"
#include "slgpu_internal.h"

namespace slgpu {

#define PH1(p, s, z, e) ((double)ph1[(p - 1) + 6 * ((s - 1) + 7 * ((z - 1) + 30 * (e - 1)))])
#define PH2(p, z, e) ((double)ph2[(p - 1) + 7 * ((z - 1) + 30 * (e - 1))])
// bf_ph1(6,..)**2, bf_ph2(5,..)**2 and bf_ph2(7,..)**2 are single-precision expressions in
// bfxsmod.f:141,148,149: the square is rounded to FP32 before it joins the FP64 sum
#define SQ32(v) ((double)((float)(v) * (float)(v)))

// bfxsmod.f:65-153: Verner 1995/96 partial photo-ionisation cross sections [Mb]
__device__ double bfxs_dev(int nz, int ne, double e, const float *__restrict__ ph1,
                           const float *__restrict__ ph2)
{
    const int ll[7] = {0, 0, 1, 0, 1, 2, 0};
    const int ninn[30] = {0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 3, 3, 3, 3, 3, 3, 3, 3,
                          5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5};
    const int ntot[30] = {1, 1, 2, 2, 3, 3, 3, 3, 3, 3, 4, 4, 5, 5, 5, 5, 5, 5,
                          6, 6, 6, 6, 6, 6, 6, 6, 6, 6, 7, 7};
    double xs = 0.0;
    if (nz < 1 || nz > 30) return xs;
    if (ne < 1 || ne > nz) return xs;
    int nout = ntot[ne - 1];
    if (nz > 18) {
        if (nz == ne) nout = 7;
        if (nz == ne + 1 && (nz == 20 || nz == 21 || nz == 22 || nz == 25 || nz == 26)) nout = 7;
    }
    if (e < PH1(1, nout, nz, ne)) return xs;
    const int nin = ninn[ne - 1];
    double einn;
    if (nz == 15 || nz == 17 || nz == 19 || (nz > 20 && nz != 26)) einn = 0.0;
    else einn = (ne < 3) ? (double)1.0e30f : PH1(1, nin, nz, ne);
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

// bf cross sections are cell-independent: xs(ig, ion) * pc_mbarn at the left group edge
// (physical_opacity.f:177-187), computed once per group grid / table upload.
__global__ void k_bfxs(Dev d, const float *__restrict__ ph1, const float *__restrict__ ph2,
                       const int *__restrict__ iz, const int *__restrict__ ii)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.ng * d.nion) return;
    int g = t / d.nion, k = t - g * d.nion;
    double wl = d.wl[g];
    double en = PC_H * PC_C / (PC_EV * wl);
    int z = iz[k];
    double xs = bfxs_dev(z, z - ii[k] + 1, en, ph1, ph2);
    d.bfxs[t] = xs * PC_MBARN;
}

// per-cell helpers (physical_opacity.f:81-89, 98-100, 168-169, 203) and the ff gamma^2
// interpolation coordinate per (element, cell) (physical_opacity.f:226-227)
__global__ void k_cell_prep(Dev d, const double *__restrict__ natom, const double *__restrict__ vol,
                            const double *__restrict__ nelec, const double *__restrict__ oc,
                            const double *__restrict__ ginv, int flags, double *rgg_out)
{
    const int nc = d.nx;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int nk = d.nion > d.nel ? d.nion : d.nel;
    if (nk < 1) nk = 1;
    if (t >= nc * nk) return;
    int k = t / nc, i = t - k * nc;
    double hckt = PC_H * PC_C / (PC_KB * d.temp[i]);
    if (k == 0) {
        constexpr double cthomson = 8.0 * PC_PI * (PC_E * PC_E * PC_E * PC_E) /
                                    (3.0 * (PC_ME * PC_ME) * (PC_C * PC_C * PC_C * PC_C));
        d.hckt[i] = hckt;
        d.tempinv[i] = 1.0 / d.temp[i];
        double s = (flags & SLGPU_NOTHOMSON) ? 0.0 : cthomson * nelec[i] * natom[i] / vol[i];
        if (d.mass[i] <= 0.0) s = 0.0;
        d.sig[i] = s;
        double nv = natom[i] / vol[i];
        d.hlparr[i] = nv * nv * nelec[i];
    }
    if (k < d.nion) {
        double o = oc[(size_t)i * d.nion + k];
        d.ocT[(size_t)k * nc + i] = o;
        d.ocg[(size_t)k * nc + i] = o * ginv[(size_t)i * d.nion + k];
        d.ginvT[(size_t)k * nc + i] = ginv[(size_t)i * d.nion + k];
    }
    if (k < d.nel) {
        int z = d.el_z[k];
        double gg = (double)(z * z) * PC_RYDBERG * hckt;
        rgg_out[(size_t)k * nc + i] = 5.0 * (log10(gg) + 4.0) + 1.0;
    }
}

// The cell's free-free row (physical_opacity.f:228-243): for each u node j,
// ffB[ic][j] = sum over the elements with gamma^2 coordinate rgg <= FF_NGG of
// z^2 natomfr * ((1 - dgg) g_ff[igg-1][j] + dgg g_ff[igg][j]), and ffC[ic][j] the same sum
// of z^2 natomfr * (g_ff[igg][j] - g_ff[igg-1][j]) for the backward; elements past the table
// go to the cell's ffclamp mask. One block of 128 threads per cell, thread j < FF_NU.
__global__ __launch_bounds__(128) void k_ff_table(Dev d, const double *__restrict__ rgg_tab)
{
    const int nc = d.nx, ic = blockIdx.x, j = threadIdx.x;
    if (j >= FF_NU) return;
    double b = 0.0, cs = 0.0;
    uint32_t m = 0u;
    for (int k = 0; k < d.nel; ++k) {
        const double rgg = rgg_tab[(size_t)k * nc + ic];
        if (rgg > FF_NGG) { m |= 1u << k; continue; }
        const int z = d.el_z[k];
        int igg = (int)floor(rgg);
        igg = igg < 1 ? 1 : (igg > FF_NGG - 1 ? FF_NGG - 1 : igg);
        const double dgg = rgg - igg;
        const double w = (double)(z * z) * d.natomfr[(size_t)ic * 30 + (z - 1)];
        const double lo = d.gff[(igg - 1) * FF_NU + j], hi = d.gff[igg * FF_NU + j];
        b += w * ((1.0 - dgg) * lo + dgg * hi);
        cs += w * (hi - lo);
    }
    d.ffB[(size_t)j * nc + ic] = b;
    d.ffC[(size_t)j * nc + ic] = cs;
    if (j == 0) d.ffclamp[ic] = m;
}

// Fused bb + bf + ff for one (cell, group); result into capT[g*nc + i] (cell-fastest).
// blockDim = 256 = 64 cells x 4 groups; grid = (ceil(nc/64), ceil(ng/4)).
// S: the block's 64 cells' line populations oc*ginv of every ion are staged in LDS once
// (FP32 in the precision form, whose line terms are FP32), so each line costs one LDS read
// per lane instead of a global gather of the 64 cells' populations.  A population that is
// not positive is staged as 0: its term is then exactly 0, as the reference's skip of
// such ions gives, and the line loop needs no test.
template <bool M, bool S>
__global__ __launch_bounds__(256) void k_opacity(Dev d, int flags, double ffc1,
                                                  const double *__restrict__ rgg_tab)
{
    using PT = typename std::conditional<M, float, double>::type;
    extern __shared__ __align__(16) unsigned char s_raw[];
    PT *s_oc = reinterpret_cast<PT *>(s_raw);
    // S: each wave also copies its group's line records into LDS 64 at a time (one
    // coalesced vector load), so the line loop holds only LDS reads, which complete in
    // order and pipeline, instead of a scalar load it must wait for on every line
    __shared__ LineRec s_lines[4][64];
    const int nc = d.nx, ng = d.ng;
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 64 + lane;
    if constexpr (S) {
        if (!(flags & SLGPU_NOBB)) {
            const int c0 = blockIdx.x * 64;
            for (int k = threadIdx.x; k < d.nion * 64; k += 256) {
                const int cc = c0 + (k & 63);
                const double o = cc < nc ? d.ocg[(size_t)(k >> 6) * nc + cc] : 0.0;
                s_oc[k] = o > 0.0 ? (PT)o : (PT)0;
            }
            __syncthreads();
        }
    }
    // wave-uniform group index (1-based)
    const int g = __builtin_amdgcn_readfirstlane(blockIdx.y * 4 + (threadIdx.x >> 6)) + 1;
    if (g > ng) return;
    const bool act = i < nc;
    const int ic = act ? i : nc - 1;
    const bool massive = act && d.mass[ic] > 0.0;
    const double hckt = d.hckt[ic];
    const double wlo = d.wl[g - 1], whi = d.wl[g];
    double acc = 0.0;

    if (!(flags & SLGPU_NOBB)) {
        // physical_opacity.f:128-137: subgroup-centre profile and stimulated emission
        const double wlc = 0.5 * (wlo + whi);
        const double wlinv = 1.0 / wlc;
        const double dwl = whi - wlo;
        const double expfac = 1.0 - exp(-hckt * wlinv);
        const double phi = wlc * wlc / (dwl * PC_C);
        const int l0 = d.gstart[g], l1 = d.gstart[g + 1];
        const double *ocg = d.ocg + ic;
        if constexpr (M) {
            // Precision form: FP32 terms gxs * oc * 2^(-chilw * hl), hl = hckt * log2(e) rounded
            // once from FP64 (one multiply and one v_exp_f32 per line). The terms are all
            // positive; each run of one ion's lines (at most 64, within a chunk) is summed in
            // two FP32 partials, multiplied by the population and added in FP64. phi and
            // expfac multiply the FP64 sum once.
            const float hl32 = (float)(hckt * LOG2E);
            double sum64 = 0.0;
            if constexpr (S) {
                // the group's lines in ion order (Dev::lines_ion, whose pad holds the lines left
                // in the record's ion run): per run one population read, per line
                // gxs * 2^(-chilw * hl) into two FP32 run sums of at most 32 terms each
                LineRec *ln = s_lines[threadIdx.x >> 6];
                const float *oc_lane = s_oc + lane;
                for (int base = l0; base < l1; base += 64) {
                    const int n = l1 - base < 64 ? l1 - base : 64;
                    wave_lds_sync();
                    if (lane < n) ln[lane] = d.lines_ion[base + lane];
                    wave_lds_sync();
                    for (int j = 0; j < n;) {
                        const int ion = __builtin_amdgcn_readfirstlane(ln[j].ion);
                        int len = __builtin_amdgcn_readfirstlane(ln[j].pad);
                        len = len < n - j ? len : n - j;
                        len = len > 1 ? len : 1;                     // always advances
                        const float of = oc_lane[ion * 64];
                        float r0 = 0.0f, r1 = 0.0f;
                        int t = j;
                        for (; t + 2 <= j + len; t += 2) {
                            r0 = fmaf(ln[t].gxs, __builtin_amdgcn_exp2f(-(ln[t].chilw * hl32)), r0);
                            r1 = fmaf(ln[t + 1].gxs, __builtin_amdgcn_exp2f(-(ln[t + 1].chilw * hl32)), r1);
                        }
                        if (t < j + len) r0 = fmaf(ln[t].gxs, __builtin_amdgcn_exp2f(-(ln[t].chilw * hl32)), r0);
                        sum64 += (double)((r0 + r1) * of);
                        j += len;
                    }
                }
            } else {
                for (int base = l0; base < l1; base += 64) {
                    const int lend = l1 - base < 64 ? l1 : base + 64;
                    float s0 = 0.0f;
                    for (int l = base; l < lend; ++l) {
                        const LineRec r = d.lines[l];
                        // (float)o > 0 exactly when o > 0 gives a nonzero term
                        const float of = (float)ocg[(size_t)r.ion * nc];
                        const float t = r.gxs * of * __builtin_amdgcn_exp2f(-(r.chilw * hl32));
                        s0 += (of > 0.0f) ? t : 0.0f;
                    }
                    sum64 += (double)s0;
                }
            }
            acc += sum64 * phi * expfac;
        } else {
            if constexpr (S) {
                LineRec *ln = s_lines[threadIdx.x >> 6];
                for (int base = l0; base < l1; base += 64) {
                    const int n = l1 - base < 64 ? l1 - base : 64;
                    wave_lds_sync();
                    if (lane < n) ln[lane] = d.lines[base + lane];
                    wave_lds_sync();
                    #pragma unroll 2
                    for (int j = 0; j < n; ++j) {
                        const LineRec r = ln[j];
                        const double o = s_oc[r.ion * 64 + lane];
                        // the select keeps the product and the sum separately rounded
                        const double term = phi * (double)r.gxs * o * exp(-((double)r.chilw * hckt)) * expfac;
                        acc += (o > 0.0) ? term : 0.0;
                    }
                }
            } else {
                for (int l = l0; l < l1; ++l) {
                    const LineRec r = d.lines[l];         // uniform: scalar loads
                    const double o = ocg[(size_t)r.ion * nc];
                    const double term = phi * (double)r.gxs * o * exp(-((double)r.chilw * hckt)) * expfac;
                    acc += (o > 0.0) ? term : 0.0;
                }
            }
        }
    }
    if (!(flags & SLGPU_NOBF)) {
        // physical_opacity.f:177-189 (cross sections precomputed per group and ion)
        const double *xs = d.bfxs + (size_t)(g - 1) * d.nion;
        for (int k = 0; k < d.nion; ++k) {
            const double x = xs[k];
            if (x != 0.0) acc += x * d.ocT[(size_t)k * nc + ic];
        }
    }
    if (!(flags & SLGPU_NOFF)) {
        // physical_opacity.f:210-243 (left group edge)
        const double wl = wlo;
        const double u = hckt * (1.0 / wl);
        const double ru = 10.0 * (log10(u) + 4.0) + 1.0;
        int iu = (int)floor(ru);
        iu = iu < 1 ? 1 : (iu > FF_NU - 1 ? FF_NU - 1 : iu);
        const double du = ru - iu;
        const double help = ffc1 * sqrt(hckt) * (1.0 - exp(-u)) * (wl * wl * wl) * d.hlparr[ic];
        acc += help * ff_gaunt_sum(d, ic, iu, du, rgg_tab);
    }
    if (act) d.capT[(size_t)(g - 1) * nc + i] = massive ? acc : 0.0;
}

// capT (ng, nc) cell-fastest -> cap64 / gas_cap / gas_emiss (nc, ng) group-fastest
// (physical_opacity.f:267-279), 32x32 tiles through LDS.
__global__ __launch_bounds__(256) void k_opacity_finish(Dev d, int want64)
{
    __shared__ double tile[32][33];
    const int nc = d.nx, ng = d.ng;
    const int g0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        int g = g0 + r, i = i0 + tx;
        tile[r][tx] = (g < ng && i < nc) ? d.capT[(size_t)g * nc + i] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        int i = i0 + r, g = g0 + tx;
        if (i < nc && g < ng) {
            double c = tile[tx][r];
            size_t o = (size_t)i * ng + g;
            if (want64) d.cap64[o] = c;
            d.cap32[o] = (float)c;
            double em = 0.0;
            if (d.mass[i] > 0.0) {
                double wl = 0.5 * (d.wl[g + 1] + d.wl[g]);
                em = c * ((wl * wl / PC_C) * planck(wl, d.temp[i]));
            }
            d.emiss[o] = (float)em;
        }
    }
}


// Planck group integrals specintv(1/T) for the DDMC/emission tables, and the Planck and
// Rosseland means of the FP32 group opacity.  One 256-thread block per cell.
// One block per cell and slice of 1024 groups: specarr, and the slice's partial sums of the
// Planck and Rosseland means (fixed-order block sums) into ctpart; k_cell_tables_sum adds the
// slices in a fixed order.
__global__ __launch_bounds__(256) void k_cell_tables(Dev d)
{
    __shared__ double sh[4];
    const int i = blockIdx.x, ng = d.ng, nc = d.nx, ns = gridDim.y;
    const double tinv = d.tempinv[i];
    const double sg = d.sig[i];
    const int gend = (int)blockIdx.y * 1024 + 1024 < ng ? (int)blockIdx.y * 1024 + 1024 : ng;
    double pg = 0.0, pr = 0.0;
    for (int g = blockIdx.y * 1024 + threadIdx.x + 1; g <= gend; g += 256) {
        const double c = (double)d.cap32[(size_t)i * ng + g - 1];
        const double s = specintv_g(tinv, d.wlinv, g);
        d.specarr[(size_t)i * ng + g - 1] = s;
        pg += c * s;
        pr += (1.0 / (sg + c)) * spectintv_g(tinv, d.wlinv, g);
    }
    pg = block_sum256(pg, sh);
    pr = block_sum256(pr, sh);
    if (threadIdx.x == 0) {
        d.ctpart[((size_t)0 * ns + blockIdx.y) * nc + i] = pg;
        d.ctpart[((size_t)1 * ns + blockIdx.y) * nc + i] = pr;
    }
}

// The means from the slices' partials: one wave per cell, lane y takes slices y, y+64, ...,
// then a fixed-order wave sum.
__global__ __launch_bounds__(64) void k_cell_tables_sum(Dev d, int ns)
{
    const int i = blockIdx.x, lane = threadIdx.x, nc = d.nx;
    if (i >= nc) return;
    double pg = 0.0, pr = 0.0;
    for (int y = lane; y < ns; y += 64) {
        pg += d.ctpart[((size_t)0 * ns + y) * nc + i];
        pr += d.ctpart[((size_t)1 * ns + y) * nc + i];
    }
    pg = wave_sum(pg);
    pr = wave_sum(pr);
    if (lane == 0) {
        d.capgrey[i] = pg;
        d.capross[i] = 1.0 / pr;
    }
}

// ---------------------------------------------------------------------------
static double *g_rgg_scratch(slgpu_ctx *c)
{
    // per (element, cell) ff coordinate, lives in the ocT allocation's tail
    return c->d.ocT + (size_t)c->nion * c->nx;
}

void launch_bfxs(slgpu_ctx *c, const float *d_ph1, const float *d_ph2, const int *d_iz,
                 const int *d_ii)
{
    int n = c->ng * c->nion;
    if (n == 0) return;
    hipLaunchKernelGGL(k_bfxs, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d, d_ph1,
                       d_ph2, d_iz, d_ii);
}

void launch_cell_prep(slgpu_ctx *c, const double *d_natom, const double *d_vol,
                      const double *d_nelec, const double *d_oc, const double *d_ginv, int flags)
{
    int nk = c->nion > c->d.nel ? c->nion : c->d.nel;
    if (nk < 1) nk = 1;
    int n = c->nx * nk;
    hipLaunchKernelGGL(k_cell_prep, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->d,
                       d_natom, d_vol, d_nelec, d_oc, d_ginv, flags, g_rgg_scratch(c));
    if (!(flags & SLGPU_NOFF))
        hipLaunchKernelGGL(k_ff_table, dim3(c->nx), dim3(128), 0, c->stream, c->d,
                           (const double *)g_rgg_scratch(c));
}

// free-free prefactor of physical_opacity.f:51-52 (cgs)
double ff_prefactor()
{
    return 4.0 * (PC_E * PC_E * PC_E * PC_E * PC_E * PC_E) /
           (3.0 * PC_H * PC_ME * (PC_C * PC_C * PC_C * PC_C)) *
           std::sqrt(PC_PI2 / (3 * PC_ME * PC_H * PC_C));
}

const double *rgg_scratch(slgpu_ctx *c) { return g_rgg_scratch(c); }

void launch_opacity(slgpu_ctx *c, int flags)
{
    const double ffc1 = ff_prefactor();
    dim3 grid((c->nx + 63) / 64, (c->ng + 3) / 4);
    const double *rgg = g_rgg_scratch(c);
    const bool mixed = c->d.precision == SLGPU_MIXED;
    // staged populations: nion x 64 cells, up to 64 KiB of LDS per block beside the 4 KiB
    // of line records
    const size_t lds = (size_t)c->nion * 64 * (mixed ? sizeof(float) : sizeof(double));
    const bool staged = !c->force_unstaged && c->nion > 0 && lds <= 65536;
    if (mixed && staged)
        hipLaunchKernelGGL((k_opacity<true, true>), grid, dim3(256), lds, c->stream, c->d, flags,
                           ffc1, rgg);
    else if (mixed)
        hipLaunchKernelGGL((k_opacity<true, false>), grid, dim3(256), 0, c->stream, c->d, flags,
                           ffc1, rgg);
    else if (staged)
        hipLaunchKernelGGL((k_opacity<false, true>), grid, dim3(256), lds, c->stream, c->d, flags,
                           ffc1, rgg);
    else
        hipLaunchKernelGGL((k_opacity<false, false>), grid, dim3(256), 0, c->stream, c->d, flags,
                           ffc1, rgg);
}

void launch_opacity_finish(slgpu_ctx *c)
{
    dim3 grid((c->ng + 31) / 32, (c->nx + 31) / 32);
    hipLaunchKernelGGL(k_opacity_finish, grid, dim3(256), 0, c->stream, c->d, 1);
    launch_cell_tables(c);
}

void launch_cell_tables(slgpu_ctx *c)
{
    const int ns = (c->ng + 1023) / 1024;
    hipLaunchKernelGGL(k_cell_tables, dim3(c->nx, ns), dim3(256), 0, c->stream, c->d);
    hipLaunchKernelGGL(k_cell_tables_sum, dim3(c->nx), dim3(64), 0, c->stream, c->d, ns);
}

}  // namespace slgpu
