// Group samplers of the transport (emitgroup and the four lumped DDMC tables) and the exact
// bucket maps that turn each of them into one table load for most draws.
//
// Every sampler is a monotone step function of its uniform draw r (inverse-CDF searches:
// GRID/emitgroup.f:23-51; TRANSPORT1/diffusion11.f90:289-313, 375-398, 447-460, 489-511,
// 571-584).  Split (0,1] into BT_K buckets b = [b/BT_K, (b+1)/BT_K); scaling by a power of
// two is exact, so the bucket of a draw is floor(r*BT_K).  Evaluating the direct search at
// the two extreme doubles of a bucket bounds the result of every draw inside it; when both
// ends agree the bucket is pure and its entry holds the answer itself, otherwise it holds
// the answer range and the search only scans that range.  Results are identical to the
// direct searches for every draw.  This needs monotone rows, so every cumulative row is a
// running sum in index order (as the reference's loops accumulate); the one row that can
// hold a NaN is the effective-scatter row of a cell whose lumped emission fraction is
// capped at 1, where the scatter event itself has probability zero up to rounding.
//
// Entry layout (uint32): bit 31 pure; pure: bit 30 = destination optically thin (the
// particle leaves DDMC for IMC), bit 29 = destination group lumped (the next DDMC step
// uses the cell's lumped thresholds), bits 0-14 = group; impure: bits 15-29 = hi, 0-14 = lo
// (group numbers for emitgroup, row positions for the lumped tables).  ng <= 32767.
// The two destination bits are what the particle's next step would otherwise load the
// group opacity for, so a pure entry hands that step its thresholds' address directly.
#pragma once
#include "slgpu_device.h"
#include "slgpu_internal.h"

namespace slgpu {

constexpr int BT_BITS = 14;
constexpr int BT_K = 1 << BT_BITS;
constexpr uint32_t BT_PURE = 0x80000000u, BT_THIN = 0x40000000u, BT_LUMP = 0x20000000u;
enum { TAB_L = 0, TAB_R = 1, TAB_D = 2, TAB_S = 3, TAB_E = 4, NTAB = 5 };

__device__ __forceinline__ int bt_key(double r)
{
    const int k = (int)(r * (double)BT_K);
    return k > BT_K - 1 ? BT_K - 1 : (k < 0 ? 0 : k);
}

// smallest and largest double with key b (the last bucket also holds r == 1)
__device__ __forceinline__ double bt_lo(int b) { return (double)b * (1.0 / BT_K); }
__device__ __forceinline__ double bt_hi(int b)
{
    return b == BT_K - 1 ? 1.0 : nextafter((double)(b + 1) * (1.0 / BT_K), 0.0);
}

__device__ __forceinline__ const uint32_t *bt_row(const Dev &d, int t, int ic)
{
    return d.btab + ((size_t)t * d.nx + (ic - 1)) * BT_K;
}

__device__ __forceinline__ double cap_of(const Dev &d, int g, int ic)
{
    return (double)d.cap32[(size_t)(ic - 1) * d.ng + (g - 1)];
}

// optical depth test of particle_advance.f90:186-193 / transport11.f90:331-336 /
// diffusion11.f90:329-355: true when group g of cell lc is transported by IMC
__device__ __forceinline__ bool is_thin(const Dev &d, int g, int lc)
{
    const double dx = d.xarr[lc] - d.xarr[lc - 1];
    return d.puretran || (d.sig[lc - 1] + cap_of(d, g, lc)) * dx < d.tauddmc;
}

// BT_LUMP when a DDMC step of group g in cell lc runs lumped: glump = nlump > 0 exactly when
// cap*dx >= taulump (diffusion11.f90:156-160) and the cell has lumped groups at all
// (nlump = 0 when speclump = 0, :107-110).  opacity: the group's cap as a double.
__device__ __forceinline__ uint32_t lump_flag(const Dev &d, double opacity, double dx, int lc)
{
    return (opacity * dx >= d.taulump && d.cell[lc - 1].nlump > 0) ? BT_LUMP : 0u;
}

// BT_THIN and BT_LUMP of group g in cell lc from one opacity load
__device__ __forceinline__ uint32_t group_flags(const Dev &d, int g, int lc)
{
    const double dx = d.xarr[lc] - d.xarr[lc - 1];
    const double cp = cap_of(d, g, lc);
    if (d.puretran || (d.sig[lc - 1] + cp) * dx < d.tauddmc) return BT_THIN;
    return lump_flag(d, cp, dx, lc);
}

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

// GRID/emitgroup.f:3-54 by direct search: binsrch over the binned emission probabilities,
// then the first group of the bin whose running sum exceeds r*capgrey (r*capemitgrey in NLTE)
__device__ __forceinline__ int emitgroup_direct(const Dev &d, double r, int ic)
{
    const int ng = d.ng;
    const double *ep = d.emitprob + (size_t)(ic - 1) * d.nep;
    const double r1 = r * d.capemgrey[ic - 1];
    int iep = binsrch(r1, ep, d.nep, true);
    if (iep > d.nep - 1) iep = d.nep - 1;
    if (iep < 0) iep = 0;
    const int ig = iep * d.nepg + 1;
    const int igp1 = ig + d.nepg - 1 < ng ? ig + d.nepg - 1 : ng;
    const double *cum = d.emitcum + (size_t)(ic - 1) * ng;
    return first_above(cum, ig - 1, igp1 - 1, r1) + 1;
}

// emitgroup restricted to results in [glo, ghi]: the bin index (binsrch with widerange,
// clamped to 0..nep-1, counts the bins whose cumulative value is <= r1) and the in-bin
// position are both monotone in r1, so each search resumes at the bucket's lower answer
__device__ __forceinline__ int emitgroup_range(const Dev &d, double r, int ic, int glo, int ghi)
{
    const int ng = d.ng, nepg = d.nepg;
    const double *ep = d.emitprob + (size_t)(ic - 1) * d.nep;
    const double r1 = r * d.capemgrey[ic - 1];
    int iep = (glo - 1) / nepg;
    const int iep_hi = (ghi - 1) / nepg;
    while (iep < iep_hi && ep[iep] <= r1) ++iep;
    const int ig = iep * nepg + 1;
    const int igp1 = ig + nepg - 1 < ng ? ig + nepg - 1 : ng;
    const double *cum = d.emitcum + (size_t)(ic - 1) * ng;
    int j = (ig > glo ? ig : glo) - 1;
    const int hi = igp1 - 1;
    while (j < hi && !(cum[j] > r1)) ++j;
    return j + 1;
}

// one lumped sampling row: cum[0..cnt) with answer positions lo + j (reversed for the
// Doppler row of a contracting cell, whose reference loop runs downwards)
struct LumpRow {
    const double *cum;
    int cnt, lo;
    bool rev;
};

__device__ __forceinline__ LumpRow lump_row(const Dev &d, int t, int ic, int nlump)
{
    const size_t row = (size_t)(ic - 1) * d.ng;
    LumpRow L;
    L.lo = t == TAB_S ? nlump : 0;
    L.cnt = t == TAB_S ? d.ng - nlump : nlump;
    L.cum = (t == TAB_L ? d.cumL : (t == TAB_R ? d.cumR : (t == TAB_D ? d.cumD : d.cumS))) + row + L.lo;
    L.rev = t == TAB_D && d.divv[ic - 1] < 0;
    return L;
}

// group at scan index j in [0, cnt] (cnt: no running sum exceeded r -> last visited)
__device__ __forceinline__ int lump_group(const Dev &d, const LumpRow &L, int ic, int j)
{
    const int pos = L.rev ? (j < L.cnt ? L.cnt - 1 - j : 0) : L.lo + (j < L.cnt ? j : L.cnt - 1);
    return d.gorder[(size_t)(ic - 1) * d.ng + pos];
}

// BT_THIN / BT_LUMP of the group a table's event moves the particle to, in the cell that
// receives it (0 past the outer face: the particle escapes instead)
__device__ __forceinline__ uint32_t dest_flags(const Dev &d, int t, int ic, int iiig)
{
    int lc = ic, g = iiig;
    if (t == TAB_L) lc = ic - 1;
    else if (t == TAB_R) lc = ic + 1;
    else if (t == TAB_D) {
        if (d.divv[ic - 1] >= 0) g = iiig + 1 > d.ng ? d.ng : iiig + 1;
        else g = iiig - 1 < 1 ? 1 : iiig - 1;
    }
    if (lc < 1 || lc > d.nx) return 0u;
    return group_flags(d, g, lc);
}

// sampled group of table t in cell ic for draw r, given the draw's bucket entry e of that
// table; fl = the destination's BT_THIN / BT_LUMP bits of a pure entry, or -1 when the
// caller must evaluate them itself (impure bucket)
__device__ __forceinline__ int sample_entry(const Dev &d, int t, int ic, int nlump, double r,
                                            uint32_t e, int &fl)
{
    if (e & BT_PURE) {
        fl = (int)(e & (BT_THIN | BT_LUMP));
        return (int)(e & 0x7fffu);
    }
    fl = -1;
    const int lo = (int)(e & 0x7fffu), hi = (int)((e >> 15) & 0x7fffu);
    if (t == TAB_E) return emitgroup_range(d, r, ic, lo, hi);
    const LumpRow L = lump_row(d, t, ic, nlump);
    int j = lo;
    while (j < hi && !(L.cum[j] > r)) ++j;
    return lump_group(d, L, ic, j);
}

// the same with the entry loaded here; `thin` is -1 (impure bucket: not evaluated) or 0/1
__device__ __forceinline__ int sample_group(const Dev &d, int t, int ic, int nlump, double r,
                                            int &thin)
{
    int fl;
    const int g = sample_entry(d, t, ic, nlump, r, bt_row(d, t, ic)[bt_key(r)], fl);
    thin = fl < 0 ? -1 : ((fl & (int)BT_THIN) ? 1 : 0);
    return g;
}

}  // namespace slgpu
