// Internal state of an slgpu context and the kernel launchers shared between the
// translation units of libslgpu.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/slgpu.h"
#include "slgpu_device.h"

namespace slgpu {

// One packed line record (the bb_xs fields the binning needs), 16 bytes.
struct LineRec {
    float gxs;      // g*f*(pi e^2/m_e c)
    float chilw;    // lower-level energy [1/cm]
    int ion;        // index into the population list
    int pad;        // in the ion-ordered copy: lines left in this record's ion run
};

// Per-cell DDMC constants (diffusion11.f90:95-152), evaluated once per iteration.
struct CellDDMC {
    double emgreyinv;   // 1/capgrey (LTE) or 1/capemitgrey (NLTE), nan -> 0
    double speclump, emitlump, caplump, doplump;
    int nlump;
    int pad;
};

// Event thresholds of one DDMC step in a cell (lumped) or a (cell, group) (unlumped),
// diffusion11.f90:243-283: 1/denominator and the cumulative Doppler / left / right
// probabilities the event draw is compared against.
struct DdmcProb {
    double denom, t1, t2, t3;
};

// Everything the kernels read or write, passed to kernels by value.
struct Dev {
    int nx, ng, nep, nepg, nion, nel;
    double tauddmc, taulump;
    int puretran, precision;
    int ampfact, thm_aniso;     // !in_trn_noamp, in_thm_aniso
    int64_t max_steps;
    uint32_t seed;
    // mesh (nx+1 / nx)
    const double *xarr, *vxarr, *divv, *vol;
    const double *dvdr, *c3divv;  // nx: (v_{i+1}-v_i)/dx_i and 3c/divv_i (transport11.f90:130,145-149)
    // groups (ng+1)
    const double *wl, *wlinv;
    // lines
    const LineRec *lines;
    const LineRec *lines_ion;   // the same records, sorted by ion within each group (backward)
    const int *gstart;          // ng+2: lines of group g (1-based) are [gstart[g], gstart[g+1])
    // tables
    double *bfxs;               // ng*nion, xs(ig, k)*1e-18 (Mb -> cm^2)
    const double *gff;          // 81*41 Fortran order
    // per-iteration gas state (cell-fastest, [k*nc + i])
    double *hckt, *hlparr, *ocg, *ocT, *sig, *tempinv, *temp, *mass, *natomfr;
    double *ffB;                          // FF_NU*nx, cell fastest: per cell, the sum over the unclamped elements
                                          // of z^2 natomfr times g_ff interpolated in gamma^2, at
                                          // each u node (physical_opacity.f:228-243)
    uint32_t *ffclamp;                    // nx: bit k set when element k's gamma^2 coordinate lies
                                          // past the table, where g_ff is clamped to >= 1 per group
    double *ffC;                          // FF_NU*nx: as ffB, with the g_ff rows' difference in
                                          // gamma^2 (the backward's slope of g_ff in rgg)
    double *ctpart;                       // 2*((ng+1023)/1024)*nx: the means' per-slice partial sums
    double *gmpart;                       // 2*((ng+63)/64)*nx: the means backward's per-tile
                                          // partial sums of dL/dT and dL/dsig
    double *ffW;                          // FF_NU*nx: backward, sum over groups of G*help spread
                                          // onto the two u nodes of each group
    int *el_z;                  // nel present elements (ascending iz)
    // opacity products
    double *capT;               // ng*nc cell-fastest scratch
    double *cap64;              // nc*ng group-fastest
    float *cap32, *emiss;       // nc*ng group-fastest (gas_cap, gas_emiss)
    double *capgrey, *capross;
    // emission opacity of the DDMC prep and the emission samplers: grd_capemit and
    // grd_capemitgrey after an NLTE build (nlte = 1), cap32 and capgrey after an LTE one
    const float *capem32;
    const double *capemgrey;
    int nlte;
    // DDMC prep
    double *specarr;            // nc*ng  specintv(tempinv(ic)) group-fastest
    double *opaclump;           // 11*nc
    double *emitprob;           // nep*nc
    double *emitcum;            // nc*ng  cumulative emission probability inside each bin
    double *emit;               // nc
    double *radtemp;            // nc
    double *srccum;             // nc*ng  cumulative normalised Planck at T_rad
    CellDDMC *cell;             // nc
    DdmcProb *cprob;            // nc     lumped-group step of each cell
    DdmcProb *gprob;            // nc*ng  unlumped step of each (cell, group), group-fastest
    int *gorder;                // nc*ng  lumped groups ascending, then unlumped ascending
    double *cumL, *cumR, *cumD, *cumS;   // nc*ng cumulative lumped sampling weights
    uint32_t *btab;             // 5*nc*BT_K bucket maps of the cumL/R/D/S rows and emitgroup
    // particles (SoA)
    int64_t npart;
    const int64_t *pstart;      // nc+1
    double *px, *pmu, *pwl, *pe, *pe0;
    uint8_t *pstat;
    int *pig, *pix, *perr;
    int64_t *pnstep;
    // tallies
    // inputs of the last opacity build kept for its backward, and the backward's scratch
    double *natom_in, *vol_in, *nelec_in; // nc
    double *ginvT;                        // nion*nc (ion-major)
    double *gcapT;                        // ng*nc cell-fastest dL/dcap after the means
    double *gcell;                        // 4*nc: dL/dT (means), dL/dsig, dL/dhckt, dL/dhlparr
    double *gfr, *goc;                    // 30*nc, nion*nc element- / ion-major
    double *tally, *jrad, *tot;           // nc, jrad_copies*nc*ng, 8
    double *eamp;                         // nc grd_eamp
    int jrad_copies;                      // replicas of grd_jrad that spread atomic traffic
    unsigned long long *ct;               // 10
    unsigned long long *queue;            // work counter of the persistent transport
    double *lum, *dev;                    // ng
    int *num;                             // ng
    int *flags;                           // bad input flags raised by kernels
    // LTE EOS: the host's ion data (ionsmod.f ion_el) and the state of the last solve
    int *ion_z, *ion_i;                   // nion: the ion list of slgpu_set_ions on the device
    int *eos_ni;                          // 30
    double *eos_e, *eos_q, *eos_g1;       // 30*31 [iz-1][ii-1]
    double *eos_amass;                    // 30
    double *eos_sahac2;                   // nc: Saha constant of the last evaluation (0: not solved)
};

// NLTE hydrogen (physical_opacity_nlte.f): the species data of slgpu_set_nlte_data, the
// per-call inputs and the work arrays, passed by value to the NLTE kernels only.
struct NlteDev {
    int n = 0, nlin = 0, ncoll = 0, kH = -1;    // levels, lines, collision lines, H I ion index
    double *glev, *wl0, *f, *gl, *gu, *aul, *bul, *blu;   // nlev / nlin
    double *cc, *dele;                          // 4*ncoll, ncoll
    double *pic, *rrc, *delepi;                 // 4*nlev, 4*nlev, nlev
    int *llw, *lup, *igs;                       // nlin (igs: group of the line, 0 outside)
    int *cl, *cu, *levpi;                       // ncoll, ncoll, nlev
    int *hstart, *hidx;                         // ng+2 offsets, the lines by group in file order
    double *radtemp, *ndens;                    // nc, 2*nc (H I, H II per cell)
    double *rc, *A, *pop, *gpi;                 // nc*n*n, nc*n*n, nc*n, nc*n
    int *ipiv, *flag;                           // nc*n, 1
    double *capemitT, *emissT;                  // (ng, nc) cell-fastest
    float *capemit32;                           // (nc, ng)
    double *capemitgrey;                        // nc
};

struct Timer {
    hipEvent_t a, b;
    bool armed = false;
};

}  // namespace slgpu

struct slgpu_ctx {
    slgpu_config cfg;
    slgpu::Dev d;
    hipStream_t stream = nullptr;
    int ng = 0, nx = 0, nion = 0;
    bool have_grid = false, have_groups = false, have_lines = false, have_bf = false,
         have_ff = false, have_ions = false, have_opacity = false, have_prep = false,
         have_sources = false, have_transport = false;
    int force_unstaged = 0;                // slgpu_set_unstaged: opacity and backward unstaged
    int64_t nline = 0, nline_used = 0;
    std::vector<double> h_xarr, h_vxarr, h_divv, h_vol, h_rcell, h_wl;
    std::vector<int> h_ion_iz, h_ion_ii;
    std::vector<float> h_ph1, h_ph2;
    std::vector<float> h_wl0;
    std::vector<int16_t> h_liz, h_lii;
    std::vector<float> h_gxs, h_chilw;
    std::vector<double> h_emit_cache;
    int opac_flags = 0;         // flags of the last opacity build (its backward uses them)
    std::vector<void *> allocations;
    bool have_eos_data = false, have_eos = false;
    // inputs and outputs of the last slgpu_eos, kept on the device for slgpu_eos_vjp
    double *eos_temp = nullptr, *eos_mass = nullptr, *eos_massfr = nullptr, *eos_nelec = nullptr,
           *eos_natom = nullptr, *eos_natomfr = nullptr, *eos_oc = nullptr, *eos_ginv = nullptr;
    int *eos_niter = nullptr;
    slgpu::NlteDev nl;          // NLTE species data and work arrays (slgpu_set_nlte_data)
    bool have_nlte = false;
    void *scratch = nullptr;    // device address of the library's static scratch array on this context's GPU
    double t_ms[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    hipEvent_t ev[24];
};

namespace slgpu {
// launchers (each enqueues on ctx->stream)
void launch_bfxs(slgpu_ctx *c, const float *d_ph1, const float *d_ph2, const int *d_iz,
                 const int *d_ii);
void launch_cell_prep(slgpu_ctx *c, const double *d_natom, const double *d_vol,
                      const double *d_nelec, const double *d_oc, const double *d_ginv,
                      int flags);
void launch_eos_ndens(slgpu_ctx *c, const double *temp, const double *mass, const double *natom,
                      const double *natomfr, double *ndens);
void launch_nlte_rad_est(slgpu_ctx *c, double *radtemp, double *jrad);
void launch_opacity(slgpu_ctx *c, int flags);
void launch_opacity_finish(slgpu_ctx *c);
void launch_cell_tables(slgpu_ctx *c);
void launch_opacity_nlte(slgpu_ctx *c, int flags, double ffc1, const double *rgg);
void launch_nlte_means(slgpu_ctx *c);
double ff_prefactor();
const double *rgg_scratch(slgpu_ctx *c);
void launch_opacity_vjp(slgpu_ctx *c, int flags, const double *gc, const double *gp,
                        const double *gr, double *gT, double *gnat, double *gne, double *gfr,
                        double *goc, double ffc1, const double *rgg);
void launch_ddmc_tables(slgpu_ctx *c);
void launch_ddmc_probs(slgpu_ctx *c);
void launch_sources(slgpu_ctx *c, uint32_t iteration);
void launch_transport(slgpu_ctx *c, uint32_t iteration);
void launch_fluxtally(slgpu_ctx *c);
void launch_jrad_reduce(slgpu_ctx *c);
void launch_sample_groups(slgpu_ctx *c, int t, int direct, int64_t n, const double *r,
                          const int *cell, int *out);
void launch_eos(slgpu_ctx *c, const double *temp, const double *mass, const double *massfr,
                double acc, double *nelec, double *natom, double *natomfr, double *oc,
                double *ginv, int *niter);
void launch_eos_vjp(slgpu_ctx *c, const double *temp, const double *mass, const double *massfr,
                    const double *nelec, const double *natom, const double *natomfr,
                    const double *oc, const double *g_natom, const double *g_fr,
                    const double *g_ne, const double *g_oc, double *out_T, double *out_mass,
                    double *out_mf);

// physical_opacity.f:228-243: sum over elements of g_ff(u, gamma^2) z^2 natomfr for cell ic at
// the u coordinate (iu, du). The bilinear g_ff is linear in the table rows of u, so the
// elements whose gamma^2 coordinate lies inside the table are folded into the cell's row
// ffB (built each opacity build by k_ff_table), leaving two loads per (cell, group). Elements
// past the table (bit set in ffclamp) keep the per-group evaluation with g_ff >= 1.
__device__ inline double ff_gaunt_sum(const Dev &d, int ic, int iu, double du,
                                      const double *__restrict__ rgg_tab)
{
    const int nc = d.nx;
    const double *B = d.ffB + ic;                 // node-major: B[j * nc]
    double s = (1.0 - du) * B[(size_t)(iu - 1) * nc] + du * B[(size_t)iu * nc];
    for (uint32_t m = d.ffclamp[ic]; m != 0u; m &= m - 1u) {
        const int k = __builtin_ctz(m);
        const int z = d.el_z[k];
        const double rgg = rgg_tab[(size_t)k * nc + ic];
        int igg = (int)floor(rgg);
        igg = igg < 1 ? 1 : (igg > FF_NGG - 1 ? FF_NGG - 1 : igg);
        const double dgg = rgg - igg;
        const double *T = d.gff;
        double gff = (1.0 - du) * (1.0 - dgg) * T[(igg - 1) * FF_NU + (iu - 1)] +
                     du * (1.0 - dgg) * T[(igg - 1) * FF_NU + iu] +
                     (1.0 - du) * dgg * T[igg * FF_NU + (iu - 1)] +
                     du * dgg * T[igg * FF_NU + iu];
        gff = gff > 1.0 ? gff : 1.0;
        s += gff * (double)(z * z) * d.natomfr[(size_t)ic * 30 + (z - 1)];
    }
    return s;
}

}  // namespace slgpu
