// Backward of the opacity build (differentiation form) on gfx950: the vector-Jacobian
// product of the FP64 group opacity cap(g,i) (GAS/physical_opacity.f:93-247), the Planck
// mean (GAS/opacity_planckmean.f:18-24) and the Rosseland mean (GAS/opacity_rossmean.f:18-21)
// with respect to the build's own inputs: temperature, natom, n_e, the mass fractions and
// the ion populations oc (ginv, V and the masses are constants of the build).
//
//   k_vjp_means     16 cells x 64 groups per block: the means' gradients folded into
//                   dL/dcap, plus their direct T and sigma_Thomson terms as per-tile partials
//   k_vjp_means_sum one wave per cell: the tiles' partials of those terms
//   k_vjp_ff_fr     one thread per (element, cell): the ff mass-fraction gradient of the
//                   elements inside the g_ff table, from the per-cell node sums
//   k_vjp_opacity   one thread per (cell, group), 64 cells x 16 groups per block: the bb
//                   line terms recomputed (not stored) and ff, with per-ion and
//                   per-element accumulators in LDS flushed once per block; bb in
//                   512-thread blocks and ff in a launch of its own when the staged
//                   populations fit.  The bb loop
//                   walks a copy of the line list ordered by ion within each group, so
//                   each run of one ion's lines reads its population once and adds to its
//                   accumulator once
//   k_vjp_bf_mfma   the bf part, a contraction over groups of the cross sections with
//                   dL/dcap, on the FP64 matrix cores
//   k_vjp_finish    one thread per cell: chain rule through hckt = hc/kT, the ff prefactor
//                   (natom/V)^2 n_e and sigma_T = c_T n_e natom/V
//
// The means are taken through the FP32-rounded cap as the forward evaluates them, with the
// rounding treated as the identity.  In the precision form (SLGPU_MIXED) the recomputed
// bb line terms are FP32 in short partial sums folded into FP64, like that form's forward; bf, ff,
// the means and every accumulator stay FP64.  Elements absent from every cell are skipped
// by the forward's ff sum, so their mass-fraction gradient is returned as 0.
#include <type_traits>
#include "slgpu_device.h"
#include "slgpu_internal.h"

namespace slgpu {

constexpr double LN10 = 2.302585092994045684;
constexpr int VJP_GROUPS = 16;          // groups per block (4 per wave)

// d/dx of x^3/(e^x-1) and of x^4 e^x/(e^x-1)^2, zero where the forward caps x at 300
__device__ __forceinline__ double d_fplanck(double x)
{
    if (x >= 300.0) return 0.0;
    const double e = exp(x), em = e - 1.0;
    return (3.0 * x * x * em - x * x * x * e) / (em * em);
}

__device__ __forceinline__ double d_fdplanck(double x)
{
    if (x >= 300.0) return 0.0;
    const double e = exp(x), em = e - 1.0;
    const double f = (x * x * x * x * e) / (em * em);
    return f * (4.0 / x + 1.0 - 2.0 * e / em);
}

// d specintv_g / dT and d spectintv_g / dT (x = hc/(kT lambda) scales as 1/T)
__device__ __forceinline__ double d_specintv_dT(double tinv, const double *__restrict__ wlinv, int ig)
{
    const double a = HCK * tinv;
    const double x0 = a * wlinv[ig - 1], x1 = a * wlinv[ig];
    const double dx = fabs(x1 - x0), m = 0.5 * (x1 + x0);
    return -(FTPI4 * dx * tinv) * (fplanck(m) + m * d_fplanck(m));
}

__device__ __forceinline__ double d_spectintv_dT(double tinv, const double *__restrict__ wlinv, int ig)
{
    const double a = HCK * tinv;
    const double x0 = a * wlinv[ig - 1], x1 = a * wlinv[ig];
    const double dx = fabs(x1 - x0), m = 0.5 * (x1 + x0);
    const double v = FTPI4T * dx * fdplanck(m);
    if (v != v) return 0.0;                 // the forward's NaN guard makes this group 0
    const double r = -(FTPI4T * dx * tinv) * (fdplanck(m) + m * d_fdplanck(m));
    return r != r ? 0.0 : r;
}

// One tile of 16 cells x 64 groups per block. Each wave takes 4 of the tile's cells in
// turn, its lanes 64 consecutive groups, so cap32 and gc are read along the group axis; the
// dL/dcap values go through an LDS tile and are stored along the cell axis of gcapT. The
// per-cell sums of dL/dT and dL/dsig are wave sums in a fixed order, one partial per tile
// (gmpart), added in tile order by k_vjp_means_sum.
__global__ __launch_bounds__(256) void k_vjp_means(Dev d, const double *__restrict__ gc,
                                                   const double *__restrict__ gp,
                                                   const double *__restrict__ gr)
{
    __shared__ double s_o[64][17];
    const int ng = d.ng, nc = d.nx;
    const int c0 = blockIdx.x * 16, g0 = blockIdx.y * 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = g0 + lane + 1;                       // this lane's group (1-based)
    const bool gok = g <= ng;
    const int nty = gridDim.y;
    for (int cl = wave * 4; cl < wave * 4 + 4; ++cl) {
        const int i = c0 + cl;
        if (i >= nc) break;                            // wave-uniform
        const double tinv = d.tempinv[i], sg = d.sig[i], R = d.capross[i];
        const double wp = gp ? gp[i] : 0.0, wr = gr ? gr[i] : 0.0;
        const bool massive = d.mass[i] > 0.0;
        double dT = 0.0, dsig = 0.0, G = 0.0;
        if (gok) {
            const double c = (double)d.cap32[(size_t)i * ng + g - 1];
            const double s = specintv_g(tinv, d.wlinv, g);
            const double t = spectintv_g(tinv, d.wlinv, g);
            const double w = sg + c;
            G = gc ? gc[(size_t)i * ng + g - 1] : 0.0;
            G += wp * s;
            dT += wp * c * d_specintv_dT(tinv, d.wlinv, g);
            if (w > 0.0 && R != 0.0) {
                // capross = 1/sum_g t_g/(sig + cap_g)
                G += wr * R * R * t / (w * w);
                dsig += wr * R * R * t / (w * w);
                dT -= wr * R * R * d_spectintv_dT(tinv, d.wlinv, g) / w;
            }
            if (!massive) G = 0.0;
        }
        s_o[lane][cl] = G;
        dT = wave_sum(dT);
        dsig = wave_sum(dsig);
        if (lane == 0) {
            d.gmpart[((size_t)0 * nty + blockIdx.y) * nc + i] = dT;
            d.gmpart[((size_t)1 * nty + blockIdx.y) * nc + i] = dsig;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 64 * 16; k += 256) {
        const int gl = k >> 4, cl = k & 15;
        if (g0 + gl < ng && c0 + cl < nc) d.gcapT[(size_t)(g0 + gl) * nc + c0 + cl] = s_o[gl][cl];
    }
}

// dL/dT and dL/dsig of the means: one wave per cell adds the tiles' partials (lane y takes
// tiles y, y+64, ...; then a fixed-order wave sum); clears the slots the line and ff kernels
// add into
__global__ __launch_bounds__(64) void k_vjp_means_sum(Dev d, int nty)
{
    const int i = blockIdx.x, lane = threadIdx.x, nc = d.nx;
    if (i >= nc) return;
    double dT = 0.0, dsig = 0.0;
    for (int y = lane; y < nty; y += 64) {
        dT += d.gmpart[((size_t)0 * nty + y) * nc + i];
        dsig += d.gmpart[((size_t)1 * nty + y) * nc + i];
    }
    dT = wave_sum(dT);
    dsig = wave_sum(dsig);
    if (lane == 0) {
        d.gcell[0 * nc + i] = dT;
        d.gcell[1 * nc + i] = dsig;
        d.gcell[2 * nc + i] = 0.0;
        d.gcell[3 * nc + i] = 0.0;
    }
}

// dynamic LDS: s_oc[nion][64] then s_fr[nel][64].  M: the precision form, whose bb line
// terms are recomputed in FP32 as in its forward, in partial sums folded into FP64.
// bb and ff of the backward.  LDS: per-ion sums of G*base over the block's lines
// (s_bb; multiplied by ginv at the flush, since d(oc*ginv*base)/d(oc) = ginv*base and ginv
// is constant over lines and groups), per-element ff sums (s_fr), and with S the staged
// line populations of the block's 64 cells (FP32 in the precision form) and each wave's slot
// of 64 line records.  A population that is not positive contributes nothing (the forward
// skips it); it is staged as -1 in the precision form, so that a positive one that
// underflows FP32 (+0) still contributes its derivative, and as 0 in FP64.  The bf part is k_vjp_bf_mfma.
// P selects the parts: 1 the bb lines, 2 ff, 3 both; NWV waves per block share its 64 cells'
// accumulators, each wave taking VJP_GROUPS / NWV of the block's groups.
template <bool M, bool S, int P, int NWV>
__global__ __launch_bounds__(NWV * 64)
__attribute__((amdgpu_waves_per_eu(P == 1 ? 6 : 1, 8)))
void k_vjp_opacity(Dev d, int flags, double ffc1, const double *__restrict__ rgg_tab)
{
    using PT = typename std::conditional<M, float, double>::type;
    constexpr bool BB = (P & 1) != 0, FF = (P & 2) != 0;
    constexpr int NT = NWV * 64;
    extern __shared__ double s_acc[];
    __shared__ LineRec s_lines[BB ? NWV : 1][64];
    const int nc = d.nx, ng = d.ng, nion = d.nion, nel = d.nel;
    const int nrow_b = BB ? nion : 0, nrow_f = FF ? nel : 0;
    double *s_bb = s_acc, *s_fr = s_acc + (size_t)nrow_b * 64;
    PT *s_pop = reinterpret_cast<PT *>(s_acc + (size_t)(nrow_b + nrow_f) * 64);
    for (int k = threadIdx.x; k < (nrow_b + nrow_f) * 64; k += NT) s_acc[k] = 0.0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    const bool act = i < nc;
    const int ic = act ? i : nc - 1;
    if constexpr (S && BB) {
        for (int k = threadIdx.x; k < nion * 64; k += NT) {
            const int cc = blockIdx.x * 64 + (k & 63);
            const double o = cc < nc ? d.ocg[(size_t)(k >> 6) * nc + cc] : 0.0;
            s_pop[k] = o > 0.0 ? (PT)o : (PT)(M ? -1 : 0);
        }
    }
    __syncthreads();
    const double hckt = d.hckt[ic];
    double dh_acc = 0.0, dhl_acc = 0.0;
    int w_iu = 0;                                  // u node pair of the W run (0: none)
    double w_lo = 0.0, w_hi = 0.0;
    for (int q = 0; q < VJP_GROUPS / NWV; ++q) {
        const int g = __builtin_amdgcn_readfirstlane(blockIdx.y * VJP_GROUPS + wave * (VJP_GROUPS / NWV) + q) + 1;
        if (g > ng) break;
        const double G = act ? d.gcapT[(size_t)(g - 1) * nc + i] : 0.0;
        const double wlo = d.wl[g - 1], whi = d.wl[g];
        if (BB && !(flags & SLGPU_NOBB)) {
            const double wlc = 0.5 * (wlo + whi);
            const double wlinv = 1.0 / wlc;
            const double dwl = whi - wlo;
            const double ec = exp(-hckt * wlinv);
            const double expfac = 1.0 - ec;
            const double phi = wlc * wlc / (dwl * PC_C);
            const int l0 = d.gstart[g], l1 = d.gstart[g + 1];
            const double *ocg = d.ocg + ic;
            double s0 = 0.0, s1 = 0.0;
            if constexpr (M) {
                // FP32 line terms as in the forward's precision form: base = gxs * 2^(-chilw*hl),
                // hl = hckt*log2(e) rounded once from FP64. phi * expfac multiplies the FP64
                // sums. The terms are positive, and FP32 partial sums of at most 32 lines are
                // folded into FP64.
                const float hl32 = (float)(hckt * LOG2E);
                const double pe = phi * expfac;
                double t0 = 0.0, t1 = 0.0;
                if constexpr (S) {
                    // the group's lines by ion: one population load and one accumulator add
                    // per run of an ion's lines (the run test is wave-uniform); a population
                    // that is not positive (stored as -1) contributes no term and no gradient
                    LineRec *ln = s_lines[wave];
                    int cur = -1;
                    bool pop_ok = false;
                    float ofc = 0.0f, sb = 0.0f;
                    double sbd = 0.0;
                    for (int base0 = l0; base0 < l1; base0 += 64) {
                        const int n = l1 - base0 < 64 ? l1 - base0 : 64;
                        wave_lds_sync();
                        if (lane < n) ln[lane] = d.lines_ion[base0 + lane];
                        wave_lds_sync();
                        // one run of an ion's lines at a time (pad: lines left in the run),
                        // at most 32 lines per FP32 partial, as a counted loop
                        for (int j = 0; j < n;) {
                            const int ion = __builtin_amdgcn_readfirstlane(ln[j].ion);
                            int len = __builtin_amdgcn_readfirstlane(ln[j].pad);
                            len = len < n - j ? len : n - j;
                            len = len < 32 ? len : 32;
                            len = len > 1 ? len : 1;                 // always advances
                            if (ion != cur) {
                                const double sbt = sbd + (double)sb;
                                if (pop_ok && sbt != 0.0) atomicAdd(&s_bb[cur * 64 + lane], G * pe * sbt);
                                cur = ion;
                                sb = 0.0f;
                                sbd = 0.0;
                                const float o = s_pop[ion * 64 + lane];
                                pop_ok = o >= 0.0f;
                                ofc = pop_ok ? o : 0.0f;
                            }
                            float a0 = 0.0f, a1 = 0.0f;
                            for (int t = j; t < j + len; ++t) {
                                const LineRec r = ln[t];
                                const float base = r.gxs * __builtin_amdgcn_exp2f(-(r.chilw * hl32));
                                const float term = base * ofc;
                                a0 += term;
                                a1 = fmaf(term, r.chilw, a1);
                                sb += base;
                            }
                            t0 += (double)a0;
                            t1 += (double)a1;
                            sbd += (double)sb;
                            sb = 0.0f;
                            j += len;
                        }
                    }
                    const double sbt = sbd + (double)sb;
                    if (pop_ok && sbt != 0.0) atomicAdd(&s_bb[cur * 64 + lane], G * pe * sbt);
                } else {
                    for (int l0b = l0; l0b < l1; l0b += 32) {
                        const int l1b = l1 - l0b < 32 ? l1 : l0b + 32;
                        float a0 = 0.0f, a1 = 0.0f;
                        for (int l = l0b; l < l1b; ++l) {
                            const LineRec r = d.lines[l];
                            const double o = ocg[(size_t)r.ion * nc];
                            if (o > 0.0) {
                                const float base = r.gxs * __builtin_amdgcn_exp2f(-(r.chilw * hl32));
                                const float term = base * (float)o;
                                a0 += term;
                                a1 = fmaf(term, r.chilw, a1);
                                atomicAdd(&s_bb[r.ion * 64 + lane], G * pe * (double)base);
                            }
                        }
                        t0 += (double)a0;
                        t1 += (double)a1;
                    }
                }
                s0 = pe * t0;
                s1 = pe * t1;
            } else {
                if constexpr (S) {
                    LineRec *ln = s_lines[wave];
                    int cur = -1;
                    double oc_ = 0.0, sb = 0.0;
                    for (int base0 = l0; base0 < l1; base0 += 64) {
                        const int n = l1 - base0 < 64 ? l1 - base0 : 64;
                        wave_lds_sync();
                        if (lane < n) ln[lane] = d.lines_ion[base0 + lane];
                        wave_lds_sync();
                        // one run of an ion's lines at a time (pad: lines left in the run)
                        for (int j = 0; j < n;) {
                            const int ion = __builtin_amdgcn_readfirstlane(ln[j].ion);
                            int len = __builtin_amdgcn_readfirstlane(ln[j].pad);
                            len = len < n - j ? len : n - j;
                            len = len > 1 ? len : 1;                 // always advances
                            if (ion != cur) {
                                if (sb != 0.0) atomicAdd(&s_bb[cur * 64 + lane], G * sb);
                                cur = ion;
                                sb = 0.0;
                                oc_ = s_pop[ion * 64 + lane];
                            }
                            if (oc_ > 0.0) {
                                for (int t = j; t < j + len; ++t) {
                                    const LineRec r = ln[t];
                                    const double base = phi * (double)r.gxs * exp(-((double)r.chilw * hckt)) * expfac;
                                    const double term = base * oc_;
                                    s0 += term;
                                    s1 += term * (double)r.chilw;
                                    sb += base;
                                }
                            }
                            j += len;
                        }
                    }
                    if (sb != 0.0) atomicAdd(&s_bb[cur * 64 + lane], G * sb);
                } else {
                    for (int l = l0; l < l1; ++l) {
                        const LineRec r = d.lines[l];
                        const double o = ocg[(size_t)r.ion * nc];
                        if (o > 0.0) {
                            const double base = phi * (double)r.gxs * exp(-((double)r.chilw * hckt)) * expfac;
                            const double term = base * o;
                            s0 += term;
                            s1 += term * (double)r.chilw;
                            atomicAdd(&s_bb[r.ion * 64 + lane], G * base);
                        }
                    }
                }
            }
            dh_acc += G * (-s1 + s0 * wlinv * ec / expfac);
        }
        if (FF && !(flags & SLGPU_NOFF)) {
            const double wl = wlo;
            const double u = hckt * (1.0 / wl);
            const double ru = 10.0 * (log10(u) + 4.0) + 1.0;
            int iu = (int)floor(ru);
            iu = iu < 1 ? 1 : (iu > FF_NU - 1 ? FF_NU - 1 : iu);
            const double du = ru - iu;
            const double eu = exp(-u);
            const double hnh = ffc1 * sqrt(hckt) * (1.0 - eu) * (wl * wl * wl);   // help / hlparr
            const double help = hnh * d.hlparr[ic];
            const double dhelp = help * (0.5 / hckt + eu * (1.0 / wl) / (1.0 - eu));
            const double dru = 10.0 / (hckt * LN10), drgg = 5.0 / (hckt * LN10);
            // Elements inside the g_ff table, through the cell's rows (k_ff_table): the sum
            // fs = sum_k g_ff zz nf and its slope fsd in hckt (through ru and rgg).
            const double *B = d.ffB + ic, *C = d.ffC + ic;    // node-major, cell fastest
            const double b0 = B[(size_t)(iu - 1) * nc], b1 = B[(size_t)iu * nc];
            double fs = (1.0 - du) * b0 + du * b1;
            double fsd = (b1 - b0) * dru + ((1.0 - du) * C[(size_t)(iu - 1) * nc] + du * C[(size_t)iu * nc]) * drgg;
            // Their mass-fraction gradient, sum over groups of G help g_ff zz, is
            // zz * sum_j A_k(j) W(j), with A_k element k's g_ff row interpolated in rgg and W
            // the groups' G*help spread over their two u nodes. W accumulates in registers
            // while iu stays, and goes to ffW when it changes (k_vjp_ff_fr finishes the sum).
            if (act) {
                if (iu != w_iu) {
                    if (w_iu > 0) {
                        atomicAdd(&d.ffW[(size_t)(w_iu - 1) * nc + ic], w_lo);
                        atomicAdd(&d.ffW[(size_t)w_iu * nc + ic], w_hi);
                    }
                    w_iu = iu;
                    w_lo = 0.0;
                    w_hi = 0.0;
                }
                w_lo += G * help * (1.0 - du);
                w_hi += G * help * du;
            }
            // Elements past the table, per group, with the reference's g_ff >= 1.
            for (uint32_t m = d.ffclamp[ic]; m != 0u; m &= m - 1u) {
                const int k = __builtin_ctz(m);
                const int z = d.el_z[k];
                const double rgg = rgg_tab[(size_t)k * nc + ic];
                const int igg = FF_NGG - 1;                    // rgg > FF_NGG
                const double dgg = rgg - igg;
                const double *T = d.gff;
                const double t00 = T[(igg - 1) * FF_NU + (iu - 1)], t10 = T[(igg - 1) * FF_NU + iu];
                const double t01 = T[igg * FF_NU + (iu - 1)], t11 = T[igg * FF_NU + iu];
                double gff = (1.0 - du) * (1.0 - dgg) * t00 + du * (1.0 - dgg) * t10 +
                             (1.0 - du) * dgg * t01 + du * dgg * t11;
                double dg = ((1.0 - dgg) * (t10 - t00) + dgg * (t11 - t01)) * dru +
                            ((1.0 - du) * (t01 - t00) + du * (t11 - t10)) * drgg;
                if (!(gff > 1.0)) { gff = 1.0; dg = 0.0; }
                const double zz = (double)(z * z);
                const double nf = d.natomfr[(size_t)ic * 30 + (z - 1)];
                fs += gff * zz * nf;
                fsd += dg * zz * nf;
                atomicAdd(&s_fr[k * 64 + lane], G * help * gff * zz);
            }
            dh_acc += G * (dhelp * fs + help * fsd);
            dhl_acc += G * hnh * fs;
        }
    }
    if (FF && act && w_iu > 0) {
        atomicAdd(&d.ffW[(size_t)(w_iu - 1) * nc + ic], w_lo);
        atomicAdd(&d.ffW[(size_t)w_iu * nc + ic], w_hi);
    }
    if (act) {
        if (dh_acc != 0.0) atomicAdd(&d.gcell[2 * nc + i], dh_acc);
        if (FF) atomicAdd(&d.gcell[3 * nc + i], dhl_acc);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < (nrow_b + nrow_f) * 64; k += NT) {
        const int row = k >> 6, cell = blockIdx.x * 64 + (k & 63);
        const double v = s_acc[k];
        if (cell < nc && v != 0.0) {
            if (row < nrow_b) atomicAdd(&d.goc[(size_t)row * nc + cell], v * d.ginvT[(size_t)row * nc + cell]);
            else atomicAdd(&d.gfr[(size_t)(row - nrow_b) * nc + cell], v);
        }
    }
}

// bf part of the backward on the matrix cores: d(sum_k bfxs(g,k) oc(k,i))/d(oc(k,i)) summed
// against G(g,i), goc(k, i) += sum_g bfxs(g, k) G(g, i), is a (nion x ng) x (ng x nc) FP64
// product.  One wave per 16-ion x 16-cell tile and per slice of
// the groups (split over waves for parallelism; the partial tiles meet in FP64 atomics), four
// groups per v_mfma_f64_16x16x4f64: lane l supplies A(ion l%16, group l/16) and
// B(group l/16, cell l%16), and its accumulator r holds D(ion 4r + l/16, cell l%16) (the
// instruction's output layout, measured on gfx950).  Out-of-range ions, cells and groups
// contribute zeros.
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int BFM_SLICES = 32;
__global__ __launch_bounds__(64) void k_vjp_bf_mfma(Dev d)
{
    const int nc = d.nx, ng = d.ng, nion = d.nion;
    const int l = threadIdx.x, lr = l & 15, lk = l >> 4;
    const int n0 = blockIdx.x * 16, m0 = blockIdx.y * 16;
    const int per = ((ng + BFM_SLICES - 1) / BFM_SLICES + 3) & ~3;
    const int g0 = blockIdx.z * per, g1 = g0 + per < ng ? g0 + per : ng;
    const int ion = m0 + lr, cell = n0 + lr;
    const bool ion_ok = ion < nion, cell_ok = cell < nc;
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    for (int g = g0 + lk; g - lk < g1; g += 4) {
        const bool g_ok = g < g1;
        const double a = (ion_ok && g_ok) ? d.bfxs[(size_t)g * nion + ion] : 0.0;
        const double b = (cell_ok && g_ok) ? d.gcapT[(size_t)g * nc + cell] : 0.0;
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    #pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int k = m0 + 4 * r + lk;
        if (k < nion && cell_ok && acc[r] != 0.0) atomicAdd(&d.goc[(size_t)k * nc + cell], acc[r]);
    }
}

// chain rule to the inputs; outputs in the forward's input layouts
__global__ __launch_bounds__(256) void k_vjp_finish(Dev d, int flags, double *gT, double *gnat,
                                                    double *gne, double *gfr_out, double *goc_out)
{
    constexpr double cthomson = 8.0 * PC_PI * (PC_E * PC_E * PC_E * PC_E) /
                                (3.0 * (PC_ME * PC_ME) * (PC_C * PC_C * PC_C * PC_C));
    const int nc = d.nx;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const double T = d.temp[i], hckt = d.hckt[i];
    const double natom = d.natom_in[i], vol = d.vol_in[i], ne = d.nelec_in[i];
    const double gsig = d.gcell[1 * nc + i], ghl = d.gcell[3 * nc + i];
    const bool thomson = !(flags & SLGPU_NOTHOMSON) && d.mass[i] > 0.0;
    gT[i] = d.gcell[0 * nc + i] + d.gcell[2 * nc + i] * (-hckt / T);
    const double nv = natom / vol;
    gne[i] = ghl * nv * nv + (thomson ? gsig * cthomson * natom / vol : 0.0);
    gnat[i] = ghl * 2.0 * natom * ne / (vol * vol) + (thomson ? gsig * cthomson * ne / vol : 0.0);
    for (int z = 0; z < 30; ++z) gfr_out[(size_t)i * 30 + z] = 0.0;
    for (int k = 0; k < d.nel; ++k) gfr_out[(size_t)i * 30 + d.el_z[k] - 1] = d.gfr[(size_t)k * nc + i];
    for (int k = 0; k < d.nion; ++k) goc_out[(size_t)i * d.nion + k] = d.goc[(size_t)k * nc + i];
}

// The mass-fraction gradient of the elements inside the g_ff table (physical_opacity.f:
// 228-243): gfr[k][i] += z^2 * sum_j A_k(j) W(i, j), with A_k(j) = (1 - dgg) g_ff[igg-1][j] +
// dgg g_ff[igg][j] element k's row at the cell's rgg and W the backward's G*help node sums.
// One thread per (element, cell); elements past the table were summed per group.
__global__ __launch_bounds__(256) void k_vjp_ff_fr(Dev d, const double *__restrict__ rgg_tab)
{
    const int nc = d.nx;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.nel * nc) return;
    const int k = t / nc, i = t - k * nc;
    const double rgg = rgg_tab[(size_t)k * nc + i];
    if (rgg > FF_NGG) return;
    int igg = (int)floor(rgg);
    igg = igg < 1 ? 1 : (igg > FF_NGG - 1 ? FF_NGG - 1 : igg);
    const double dgg = rgg - igg;
    const double *lo = d.gff + (size_t)(igg - 1) * FF_NU, *hi = d.gff + (size_t)igg * FF_NU;
    const double *W = d.ffW + i;                       // node-major, cell fastest
    double s = 0.0;
    for (int j = 0; j < FF_NU; ++j) s += ((1.0 - dgg) * lo[j] + dgg * hi[j]) * W[(size_t)j * nc];
    const int z = d.el_z[k];
    d.gfr[(size_t)k * nc + i] += (double)(z * z) * s;
}

void launch_opacity_vjp(slgpu_ctx *c, int flags, const double *gc, const double *gp,
                        const double *gr, double *gT, double *gnat, double *gne, double *gfr,
                        double *goc, double ffc1, const double *rgg)
{
    const int nc = c->nx, ng = c->ng;
    (void)hipMemsetAsync(c->d.goc, 0, (size_t)std::max(c->nion, 1) * nc * 8, c->stream);
    (void)hipMemsetAsync(c->d.gfr, 0, (size_t)30 * nc * 8, c->stream);
    (void)hipMemsetAsync(c->d.ffW, 0, (size_t)FF_NU * nc * 8, c->stream);
    {
        const int nty = (ng + 63) / 64;
        hipLaunchKernelGGL(k_vjp_means, dim3((nc + 15) / 16, nty), dim3(256), 0, c->stream, c->d, gc, gp, gr);
        hipLaunchKernelGGL(k_vjp_means_sum, dim3(nc), dim3(64), 0, c->stream, c->d, nty);
    }
    const bool mixed = c->d.precision == SLGPU_MIXED;
    const size_t acc_b = (size_t)(c->nion + c->d.nel) * 64 * sizeof(double);
    const size_t bb_b = (size_t)c->nion * 64 * sizeof(double), ff_b = (size_t)c->d.nel * 64 * sizeof(double);
    const size_t pop_b = (size_t)c->nion * 64 * (mixed ? sizeof(float) : sizeof(double));
    // With the populations staged and the line slots of 8 waves within 64 KiB of LDS, the bb
    // lines run in 512-thread blocks (8 waves x 2 groups sharing the cells' accumulators) and
    // ff in a kernel of its own; otherwise both in one 256-thread kernel, populations unstaged.
    const bool staged = !c->force_unstaged && bb_b + pop_b + 8 * 64 * sizeof(LineRec) <= 65536;
    dim3 grid((nc + 63) / 64, (ng + VJP_GROUPS - 1) / VJP_GROUPS);
    if (staged) {
        if (!(flags & SLGPU_NOBB) && c->nion > 0) {
            if (mixed)
                hipLaunchKernelGGL((k_vjp_opacity<true, true, 1, 8>), grid, dim3(512), bb_b + pop_b, c->stream,
                                   c->d, flags, ffc1, rgg);
            else
                hipLaunchKernelGGL((k_vjp_opacity<false, true, 1, 8>), grid, dim3(512), bb_b + pop_b, c->stream,
                                   c->d, flags, ffc1, rgg);
        }
        if (!(flags & SLGPU_NOFF) && c->d.nel > 0) {
            if (mixed)
                hipLaunchKernelGGL((k_vjp_opacity<true, false, 2, 4>), grid, dim3(256), ff_b, c->stream,
                                   c->d, flags, ffc1, rgg);
            else
                hipLaunchKernelGGL((k_vjp_opacity<false, false, 2, 4>), grid, dim3(256), ff_b, c->stream,
                                   c->d, flags, ffc1, rgg);
        }
    } else if (mixed) {
        hipLaunchKernelGGL((k_vjp_opacity<true, false, 3, 4>), grid, dim3(256), acc_b, c->stream, c->d, flags, ffc1, rgg);
    } else {
        hipLaunchKernelGGL((k_vjp_opacity<false, false, 3, 4>), grid, dim3(256), acc_b, c->stream, c->d, flags, ffc1, rgg);
    }
    if (!(flags & SLGPU_NOFF) && c->d.nel > 0)
        hipLaunchKernelGGL(k_vjp_ff_fr, dim3((c->d.nel * nc + 255) / 256), dim3(256), 0, c->stream,
                           c->d, rgg);
    if (!(flags & SLGPU_NOBF) && c->nion > 0)
        hipLaunchKernelGGL(k_vjp_bf_mfma, dim3((nc + 15) / 16, (c->nion + 15) / 16, BFM_SLICES),
                           dim3(64), 0, c->stream, c->d);
    hipLaunchKernelGGL(k_vjp_finish, dim3((nc + 255) / 256), dim3(256), 0, c->stream, c->d,
                       flags, gT, gnat, gne, gfr, goc);
}

}  // namespace slgpu
