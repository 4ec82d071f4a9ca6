/*
 * slgpu — MI355X (gfx950) implementation of SuperLite's per-iteration hot path,
 * behind a plain C ABI callable from Fortran through iso_c_binding.
 *
 * Each entry point replaces the body of one reference subroutine (or a fixed
 * sequence of them, in the order superlite.f90:157-219 calls them).  The host keeps
 * every module array it owns; the library keeps device-resident copies of the mesh,
 * groups, line list, opacities, DDMC tables, particles and tallies between calls.
 *
 * Conventions
 *   - every function returns 0 on success, a positive SLGPU_E* code on error;
 *     slgpu_last_error_message() gives a text description.
 *   - cells ic = 1..nx map to 0-based storage; the mesh has no void cells (ic == ix).
 *   - group-resolved arrays use the reference's memory order grd_cap(ng, nc):
 *     element (ig, ic) is at [(ic-1)*ng + (ig-1)].
 *   - output pointers are host pointers and may be NULL (the result then stays on
 *     the device for the next call); input pointers are host pointers, borrowed for
 *     the duration of the call only.
 *   - one context per process and GPU; calls on one context are host-serialised and
 *     stream-ordered on the context's HIP stream; a context is not thread-safe.
 */
#ifndef SLGPU_H
#define SLGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct slgpu_ctx slgpu_ctx;

enum {
    SLGPU_OK = 0,
    SLGPU_EINVAL = 1,      /* bad argument / size */
    SLGPU_EORDER = 2,      /* call made before its inputs were set */
    SLGPU_EHIP = 3,        /* HIP runtime error */
    SLGPU_ETRANSPORT = 4,  /* transport error with errorfatal set (particle_advance.f90:262-266) */
    SLGPU_ENOMEM = 5
};

/* Arithmetic of a context.  SLGPU_FP64: the reference's FP64 (gas_cap is FP32 as in
 * physical_opacity.f:267).  SLGPU_MIXED, the precision form: FP32 bound-bound line
 * terms, summed in short FP32 partials folded into FP64, in the opacity build and its
 * backward; FP32 sampled path lengths,
 * event times, adiabatic factors and r^3-uniform positions in the transport.  The
 * shell-crossing quadratic, DDMC leakage thresholds, positions, energies and every tally
 * stay FP64. */
enum { SLGPU_FP64 = 0, SLGPU_MIXED = 1 };

/* opacity flags (inputparmod.f:63-66: in_nobbopac, in_nobfopac, in_noffopac, in_nothmson) */
enum { SLGPU_NOBB = 1, SLGPU_NOBF = 2, SLGPU_NOFF = 4, SLGPU_NOTHOMSON = 8 };

typedef struct {
    int nx;                /* grd_nx == grd_ncell (gridmod.f:15,33) */
    int ng;                /* grp_ng (groupmod.f:7) */
    int nep, nepg;         /* grd_nep, grd_nepg (gridmod.f:121-122) */
    double tauddmc;        /* trn_tauddmc (transportmod.f:14) */
    double taulump;        /* trn_taulump (transportmod.f:15) */
    int puretran;          /* in_puretran (inputparmod.f:56) */
    int errorfatal;        /* trn_errorfatal (transportmod.f:11) */
    int precision;         /* SLGPU_FP64 or SLGPU_MIXED */
    uint32_t seed;         /* key of the counter-based per-particle RNG (replaces randommod.f) */
    int64_t npartmax;      /* prt_npartmax (particlemod.f90:175) */
    int64_t max_steps;     /* per-particle step cap, a particle beyond it dies with error 200 (0: 2^26) */
    int ampfact;           /* 1: amplification factor on, i.e. in_trn_noamp = .false.
                              (inputparmod.f:57, transport11.f90:410-421); 0: the default */
    int thm_aniso;         /* 1: in_thm_aniso, the experimental anisotropic Thomson angle
                              (inputparmod.f:58, transport11.f90:258-285); 0: the default */
} slgpu_config;

/* Host gas state consumed by the opacity build (gasmod.f:20-34 and ion_grndlev,
 * ionsmod.f:35-40, packed ion-fastest for the ion list given to slgpu_set_ions). */
typedef struct {
    const double *temp;     /* gas_temp(nc) */
    const double *natom;    /* gas_natom(nc) */
    const double *vol;      /* gas_vol(nc) */
    const double *mass;     /* gas_mass(nc) */
    const double *nelec;    /* gas_nelec(nc) */
    const double *natomfr;  /* gas_natomfr(1:30, nc), iz-fastest */
    const double *oc;       /* ion_grndlev(iz,ic)%oc(ii)   as [(ic-1)*nion + k] */
    const double *ginv;     /* ion_grndlev(iz,ic)%ginv(ii) as [(ic-1)*nion + k] */
} slgpu_gas;

/* transportmod_init + gridmod/particle allocation (superlite.f90:100-116) */
int slgpu_create(const slgpu_config *cfg, slgpu_ctx **out);
int slgpu_destroy(slgpu_ctx *ctx);
const char *slgpu_last_error_message(void);

/* mesh: grd_xarr(nx+1), grd_vxarr(nx+1), grd_divv(nx), grd_vol(nx), grd_rcell(nx)
 * (GRID/grid_setup.f:16-51, GRID/grid_volume.f:80-101) */
int slgpu_set_grid(slgpu_ctx *ctx, const double *xarr, const double *vxarr,
                   const double *divv, const double *vol, const double *rcell);

/* groups: grp_wl(ng+1) (groupmod.f:45-51); the flux grid is the group grid
 * (in_flx_ndim(1) == 0, fluxmod.f:132-136) */
int slgpu_set_groups(slgpu_ctx *ctx, const double *grp_wl);

/* lambda-sorted line list bb_xs (bbxsmod.f:27-34, sorted by bbxsmod.f:126-179) */
int slgpu_set_lines(slgpu_ctx *ctx, int64_t nline, const float *wl0, const float *gxs,
                    const float *chilw, const int16_t *iz, const int16_t *ii);

/* Verner tables bf_ph1(6,7,30,30), bf_ph2(7,30,30) (bfxsmod.f:14) and
 * Sutherland ff_gff(81,41) (ffxsmod.f:18), Fortran order */
int slgpu_set_bf_tables(slgpu_ctx *ctx, const float *ph1, const float *ph2);
int slgpu_set_ff_table(slgpu_ctx *ctx, const double *gff);

/* ion list of the populations: (iz, ii) pairs, ascending, ii <= min(iz, ni(iz)-1)
 * (physical_opacity.f:95-101,165-171) */
int slgpu_set_ions(slgpu_ctx *ctx, int nion, const int *ion_iz, const int *ion_ii);

/* GAS/physical_opacity.f:3-357 (grp_ngs == 1) + GAS/opacity_planckmean.f:18-24 +
 * GAS/opacity_rossmean.f:18-21.  out_cap64 is the FP64 `cap` before the sngl() of
 * physical_opacity.f:267; out_cap32 is gas_cap; out_emiss is gas_emiss. */
int slgpu_opacity_build(slgpu_ctx *ctx, const slgpu_gas *gas, int flags,
                        double *out_cap64, float *out_cap32, double *out_sig,
                        double *out_capgrey, double *out_capross, float *out_emiss);

/* No counterpart in the reference: slgpu_opacity_build for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and errors as slgpu_opacity_build, but the per-call
 * staging of the ion populations lives in the library's static 4 MiB device scratch array
 * instead of a fresh allocation, so it calls no hipMalloc/hipFree.  Limit: 2*nion*nc doubles
 * must fit in 4 MiB (nion*nc <= 262144), else SLGPU_EINVAL.  The array is shared by every
 * call on one GPU and keeps the last call's contents: calls that overlap on one GPU (several
 * contexts or threads) race. */
int slgpu_opacity_build_noalloc(slgpu_ctx *ctx, const slgpu_gas *gas, int flags,
                                double *out_cap64, float *out_cap32, double *out_sig,
                                double *out_capgrey, double *out_capross, float *out_emiss);

/* mpimod_ser.f:98-112 (grd_* = gas_*) + SOURCE/sourceenergy.f90:23-50 +
 * GRID/leakage_opacity11.f90:30-171 + GRID/emission_probability.f90:39-62, plus the
 * per-cell sampling tables the transport uses in place of the reference's
 * per-event O(ng) scans.  radtemp = grd_radtemp(nc), lbol = in_L_bol. */
int slgpu_ddmc_prep(slgpu_ctx *ctx, const double *radtemp, double lbol,
                    double *out_opaclump /* 11*nc */, double *out_emitprob /* nep*nc */,
                    double *out_emit /* nc */, double *out_sthermal);

/* SOURCE/sourcenumbers.f90:26-169 (one rank, in_srcepwr = 1, no surface source),
 * SOURCE/vacancies.f90, SOURCE/interior_source.f90:57-168,
 * SOURCE/source_transformdirection.f90:23-46.  nstot = src_ns. */
int slgpu_sources(slgpu_ctx *ctx, int64_t nstot, uint32_t iteration,
                  int *out_nvol /* nc, may be NULL */, int64_t *out_nnew);

/* particle_advance.f90:65-325 with TRANSPORT1/transport11.f90 and diffusion11.f90.
 * out_eamp = grd_eamp (nc; zero unless cfg.ampfact, particle_advance.f90:216-217).
 * out_totals = [tot_eout, tot_evelo, tot_sflux, e_start, e_dead, e_source, sum(grd_eamp)];
 * e_start = tot_eout + tot_evelo + e_dead + sum(grd_eamp) up to round-off
 * out_counters = [npactive, npstepimc, npstepddmc, npstepmax, npmethswap, npflux,
 *                 ndead, nerror, nwarning, 0] (countersmod.f:12-20) */
int slgpu_transport(slgpu_ctx *ctx, uint32_t iteration, double *out_tally /* nc */,
                    double *out_jrad /* ng*nc */, double *out_eamp /* nc */,
                    double *out_totals /* 7 */, int64_t *out_counters /* 10 */);

/* fluxtally.f:25-54: flx_luminos, flx_lumdev, flx_lumnum over the group grid */
int slgpu_fluxtally(slgpu_ctx *ctx, double *out_luminos, double *out_lumdev,
                    int *out_lumnum);

/* per-particle final state of the last transport (stat 1 flux, 2 dead), host copies;
 * any pointer may be NULL.  n = number of particles of the last slgpu_sources. */
int slgpu_get_particles(slgpu_ctx *ctx, int64_t n, uint8_t *stat, double *x, double *mu,
                        double *wl, double *e, int *ig, int64_t *nstep, int *err);

/* NLTE hydrogen data (nltemod.f:88-166 read_atom_nlte + read_nlte_data.f:148-233): the H I
 * model atom as read_nlte_data stores it.  glev (nlev) statistical weights; per radiative
 * line (nlin): llw < lup (1-based levels), wl0 [Angstrom] from the level energies (the file's
 * value where they give none), f, gl, gu, aul, bul, blu; per electron-impact line (ncoll):
 * llw_coll < lup_coll, coll_c (4 fit coefficients per line), dele [eV]; per level: lev_pi
 * (the level of each PI/RR record, every level once), pi_c and rr_c (4 fit coefficients per
 * record), dele_pi [eV].  Level pairs must be distinct within each kind.  Call after
 * slgpu_set_groups and slgpu_set_ions (the group of each line and the H I ion index are
 * taken then).  superlite_amd/io.py reads the data.nlte.* files into these arrays. */
int slgpu_set_nlte_data(slgpu_ctx *ctx, int nlev, int nlin, int ncoll, const double *glev,
                        const int *llw, const int *lup, const double *wl0, const double *f,
                        const double *gl, const double *gu, const double *aul, const double *bul,
                        const double *blu, const int *llw_coll, const int *lup_coll,
                        const double *coll_c, const double *dele, const int *lev_pi,
                        const double *pi_c, const double *rr_c, const double *dele_pi);

/* GAS/physical_opacity_nlte.f:3-708 + opacity_planckmean.f + opacity_rossmean.f with H I in
 * NLTE (nlte_nelem = 1): per massive cell the H I level populations from the rate equations
 * (radiative rates with J = B_nu(T) at the group centre, electron-impact rates, PI/RR with
 * radtemp; solved by LU with partial pivoting), the H lines' NLTE absorption and emission,
 * the other ions' lines in LTE (H I lines of the context's line list are skipped), bf and ff.
 * radtemp (nc) = grd_radtemp; ndens (2*nc) = nlte_ndens(i,1:2,1), the H I and H II number
 * densities of eos_update.f:52-57.  Outputs as slgpu_opacity_build, plus out_capemit
 * (gas_capemit, ng*nc FP32), out_capemitgrey (nc) and out_npop (nlev*nc, level-fastest, the
 * normalised populations); any may be NULL.  FP64 arithmetic in both precision forms.
 * SLGPU_EINVAL where the reference stops on a singular rate matrix.  After this build the
 * DDMC prep and the transport's group samplers take the emission opacity from gas_capemit and
 * gas_capemitgrey, as the reference's in_nlte branches do (leakage_opacity11.f90:76-85,
 * emission_probability.f90:28-60, emitgroup.f:23-50, diffusion11.f90:136-182,576-581); after an
 * LTE build they take gas_cap and gas_capgrey. */
int slgpu_opacity_build_nlte(slgpu_ctx *ctx, const slgpu_gas *gas, const double *radtemp,
                             const double *ndens, int flags, double *out_cap64, float *out_cap32,
                             float *out_capemit, double *out_sig, double *out_capgrey,
                             double *out_capemitgrey, double *out_capross, float *out_emiss,
                             double *out_npop);

/* nltemod.f:231-260 nlte_rad_est, after slgpu_transport: from the transport's census energy
 * tally grd_tally, radtemp (in/out, nc) += 0.5 ((tally/(a vol))^(1/4) - radtemp), and out_jrad
 * (ng*nc group-fastest, may be NULL) = wlm^2 B(wlm, radtemp)/c at each group centre. */
int slgpu_nlte_rad_est(slgpu_ctx *ctx, double *radtemp, double *out_jrad);

/* Differentiation form: vector-Jacobian product of the last slgpu_opacity_build.
 * Backward of physical_opacity.f:93-247 (bb, bf, ff), opacity_planckmean.f:18-24 and
 * opacity_rossmean.f:18-21 with respect to the build's own inputs.  grad_cap64 (ng*nc, the
 * layout of out_cap64), grad_capgrey (nc) and grad_capross (nc) are dL/d(output); any may be
 * NULL (zero).  Outputs, any may be NULL: dL/dT (nc), dL/dnatom (nc), dL/dn_e (nc),
 * dL/dnatomfr (30*nc, the layout of slgpu_gas.natomfr; 0 for elements absent from every
 * cell) and dL/doc (nion*nc, the layout of slgpu_gas.oc).  The means go through the
 * FP32-rounded cap as in the forward, the rounding taken as the identity.  Density and
 * composition enter the opacity through the EOS populations (eos_update.f:37-45), which
 * the host computes: dL/drho and dL/dX follow by chaining these outputs through it.  Line
 * terms are recomputed, not stored.  Needs nion + (elements present) <= 128.
 * SLGPU_EORDER after an NLTE build: the backward covers the LTE build. */
int slgpu_opacity_vjp(slgpu_ctx *ctx, const double *grad_cap64, const double *grad_capgrey,
                      const double *grad_capross, double *out_grad_temp, double *out_grad_natom,
                      double *out_grad_nelec, double *out_grad_natomfr, double *out_grad_oc);

/* LTE EOS data of the host, as its ionsmod.f ion_el holds them after its first EOS call:
 * ni[30] = ion_el(iz)%ni (stages incl. the bare nucleus); e, q, g1 [30*31], element-major
 * ([(iz-1)*31 + ii-1]): ionisation energy [erg], partition function (frozen at the first
 * call, ionsmod.f:208-215) and ground-level weight glev(1) of stage ii; amass[30] =
 * elem_data(iz)%m [amu] (elemdatamod.f). */
int slgpu_set_eos_data(slgpu_ctx *ctx, const int *ni, const double *e, const double *q,
                       const double *g1, const double *amass);

/* GAS/gas_setup.f:101-160 (massfr2natomfr, no container fraction) + GAS/eos_update.f:23-45 +
 * ionsmod.f:160-356 (ions_solve_eos, LTE) on the device, for every cell with mass > 0.
 * temp, mass (nc); massfr (30*nc, iz-fastest); nelec (nc, in/out: the start value of the
 * secant solve, 1.0 at the first call as gasmod.f:101, then the previous solution).  acc:
 * relative accuracy of nelec (<= 0: the reference's 1e-8).  Outputs (any may be NULL):
 * natom (nc), natomfr (30*nc), oc/ginv (nion*nc, the layout of slgpu_gas.oc for the ion
 * list of slgpu_set_ions), niter (nc; 41 = accuracy not reached, the reference warns).
 * SLGPU_EINVAL where the reference stops (ionsmod.f:291,338-349; gas_setup.f:120-123). */
int slgpu_eos(slgpu_ctx *ctx, const double *temp, const double *mass, const double *massfr,
              double acc, double *nelec, double *out_natom, double *out_natomfr,
              double *out_oc, double *out_ginv, int *out_niter);

/* Device-pointer forms of the differentiation form, for callers whose arrays live on the
 * context's GPU (e.g. torch tensors on that device), so the nc x ng arrays cross no PCIe:
 * slgpu_opacity_outputs_dev copies the last build's cap64 (ng*nc), capgrey and capross (nc)
 * into device arrays (any may be NULL); slgpu_opacity_vjp_dev is slgpu_opacity_vjp with the
 * three input gradients as device arrays (outputs on the host as there).  The library works
 * on its own stream: the caller finishes writing the input gradients first, and each call
 * returns with its copies complete.  No counterpart in the reference (like the backward). */
int slgpu_opacity_outputs_dev(slgpu_ctx *ctx, double *d_cap64, double *d_capgrey, double *d_capross);
int slgpu_opacity_vjp_dev(slgpu_ctx *ctx, const double *d_grad_cap64, const double *d_grad_capgrey,
                          const double *d_grad_capross, double *out_grad_temp, double *out_grad_natom,
                          double *out_grad_nelec, double *out_grad_natomfr, double *out_grad_oc);

/* eos_update.f:47-58 with in_nlte (nlte_nelem = 1): nlte_ndens(i,1:2,1) of the last slgpu_eos,
 * the H I and H II number densities n_ii natom natomfr(1)/vol from its last Saha evaluation.
 * out_ndens (2*nc, H I then H II per cell; 0 in massless cells) is the ndens argument of
 * slgpu_opacity_build_nlte. */
int slgpu_eos_nlte_ndens(slgpu_ctx *ctx, double *out_ndens);

/* Differentiation form: vector-Jacobian product of the last slgpu_eos at its solution (the
 * implicit function theorem on the converged nelec, partition functions held fixed).
 * dL/d(natom (nc), natomfr (30*nc), nelec (nc), oc (nion*nc)) in, any may be NULL (zero);
 * out (any may be NULL): the EOS's part of dL/dT (nc), dL/dmass (nc; dL/drho = vol*dL/dmass),
 * dL/dmassfr (30*nc).  Chained after slgpu_opacity_vjp this gives the opacity's gradients
 * with respect to temperature, density and mass fractions. */
int slgpu_eos_vjp(slgpu_ctx *ctx, const double *grad_natom, const double *grad_natomfr,
                  const double *grad_nelec, const double *grad_oc, double *out_grad_temp,
                  double *out_grad_mass, double *out_grad_massfr);

/* Verification hook, no counterpart in the reference: on = 1 makes the context's opacity
 * build and its backward take the kernels they use when the ion populations do not fit in
 * LDS (more than 256 ions for the build in the precision form, 128 in FP64; about 74 and 56
 * for the backward), so tests can compare them with the staged kernels on small inputs.
 * on = 0 (the default) restores the choice by size. */
int slgpu_set_unstaged(slgpu_ctx *ctx, int on);

/* Verification hook for the transport's group samplers; no single reference routine
 * corresponds.  It evaluates, for n (draw r, cell) pairs, the sampler the transport
 * uses: table 0-3 = the lumped loops of diffusion11.f90 (left leak :375-398, right leak
 * :447-460/489-511, Doppler :289-313, effective scatter :571-584), table 4 = emitgroup.f:23-51.
 * direct = 0 goes through the bucket maps the kernels use, direct = 1 through the plain
 * searches they replace; both must agree for every r.  Requires slgpu_ddmc_prep.
 * Lumped tables of cells without lumped groups return 0.  Host arrays. */
int slgpu_sample_groups(slgpu_ctx *ctx, int table, int direct, int64_t n, const double *r,
                        const int *cell, int *out_group);

/* accumulated device time per phase in ms, measured with HIP events on the context
 * stream: [0] opacity, [1] ddmc prep, [2] sources, [3] transport, [4] fluxtally,
 * [5] opacity bb+bf+ff kernel alone, [6] transport kernel alone, [7] opacity backward,
 * [8] EOS, [9] EOS backward, [10] opacity backward kernels alone, [11] opacity forward
 * kernels alone (cell prep, bb+bf+ff, transpose, means); n <= 12; reset = 1 zeroes them */
int slgpu_timers(slgpu_ctx *ctx, double *ms, int n, int reset);

/* the HIP stream all work of the context is ordered on (hipStream_t) */
void *slgpu_stream(slgpu_ctx *ctx);

/* Measurement hook, no reference counterpart: measured throughput of the current device's
 * raw FP32 exponential (v_exp_f32) and of the FP64 exp() sequence, in evaluations per
 * second.  These are the compute ceilings of the opacity build's line loop (one
 * exponential per line and cell).  Either pointer may be NULL. */
int slgpu_peak_rates(double *out_exp32_per_s, double *out_exp64_per_s);

#ifdef __cplusplus
}
#endif
#endif
