/* ORACLE — CPU FP64 restatement of SuperLite's per-iteration hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Imported, linked or executed only by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg, always as the checker or
 * the reported CPU baseline, never as the product path.
 *
 * Each function cites the reference file:line it restates.  Indices in the model
 * struct follow the reference: cells ix = 1..nx (compressed cell ic == ix, no void
 * cells), groups ig = 1..ng; arrays are stored 0-based with the reference's memory
 * order (grd_cap(ng, nc) group-fastest: cap[(ic-1)*ng + ig-1]).
 *
 * DDMC per-cell caches (TRANSPORT1/diffusion11.f90:95-152, 275-280, 565-569) are
 * evaluated exactly for the particle's current cell on every event ("perfect cache"):
 * the Planck group integrals specarr(ig) always hold specintv(tempinv(ic)), which is
 * what the reference's cache holds once filled.  This removes the reference's
 * dependence on thread scheduling and stale entries (the reference reads a
 * not-yet-filled cache in the unlumped Doppler rate, diffusion11.f90:185-190).
 */
#ifndef SL_ORACLE_H
#define SL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    /* grid (gridmod.f:19-23, grid_setup.f:46-51, grid_volume.f:82-87) */
    int nx;
    const double *xarr, *vxarr;   /* nx+1 */
    const double *divv, *vol;     /* nx */
    /* groups (groupmod.f:45-51) */
    int ng;
    const double *wl, *wlinv;     /* ng+1 */
    int nep, nepg;                /* gridmod.f:121-122 */
    /* transport parameters (inputparmod.f:52-57) */
    double tauddmc, taulump;
    int puretran;
    int max_steps;                /* per-particle step cap (0 = none) */
    int ampfact;                  /* !in_trn_noamp: amplification factor (transport11.f90:410-421) */
    int thm_aniso;                /* in_thm_aniso: anisotropic Thomson angle (transport11.f90:258-285) */
    /* per-iteration state (grd_* after bcast_nonpermanent, mpimod_ser.f:98-112) */
    const float *cap;             /* ng*nc */
    const double *sig, *capgrey, *tempinv;   /* nc */
    const double *opaclump;       /* 11*nc */
    const double *emitprob;       /* nep*nc */
    const double *emit;           /* nc, after sourceenergy */
    const double *radtemp;        /* nc */
    const int *nvol;              /* nc, after sourcenumbers */
    /* in_nlte: the emission opacity grd_capemit (ng*nc) and grd_capemitgrey (nc) that the
     * NLTE branches of leakage_opacity11, emission_probability, emitgroup and diffusion11
     * read; both NULL for LTE */
    const float *capemit;
    const double *capemitgrey;
} or_model;

typedef struct {
    double *tally;    /* nc  grd_tally */
    double *jrad;     /* ng*nc grd_jrad */
    double *eamp;     /* nc  grd_eamp (particle_advance.f90:216-217), may be NULL */
    double *luminos, *lumdev;   /* ng (flux grid == group grid) */
    int *lumnum;
    double tot[7];    /* eout, evelo, sflux, estart(lab, after the IMC lab transform), edead, esrc(comoving),
                         eamp (sum of grd_eamp) */
    int64_t ct[10];   /* nactive, nstepimc, nstepddmc, nstepmax, nmethswap, nflux, ndead, nerr, nwarn,
                        Newton-Raphson Doppler distances (transport11.f90:159-199) */
    /* optional per-particle final state (NULL to skip) */
    uint8_t *pstat;   /* 1 flux, 2 dead */
    double *px, *pmu, *pwl, *pe;
    int *pig, *pix;
    int64_t *pnstep;
    int *perr;
} or_result;

/* ---- building blocks --------------------------------------------------- */
int or_binsrch(double x, const double *arr, int n, int widerange);          /* MISC/binsrch.f90:3-77 */
void or_specintv(double tempinv, const double *wlinv, int n, int offset, double *ss);  /* groupmod.f:146-272 mode 0 */
void or_spectintv(double tempinv, const double *wlinv, int n, double *ss);  /* groupmod.f:275-404 mode 0 */
double or_specint0(double tempinv, const double *wlinv, int ig);            /* groupmod.f:99-120 */
double or_dopspeccalc(double tempinv, const double *wlinv, int ig);         /* groupmod.f:123-143 */
double or_bfxs(int nz, int ne, double e, const float *ph1, const float *ph2); /* bfxsmod.f:65-153 */
double or_planck(double wl, double temp);                                   /* physconstmod.f:79-93 */

/* ---- opacity build: physical_opacity.f:3-357 + opacity_planckmean.f + opacity_rossmean.f ---- */
typedef struct {
    int nc, ng;
    const double *wl;                 /* ng+1 grp_wl, cm */
    const double *temp, *natom, *vol, *mass, *nelec;   /* nc */
    const double *natomfr;            /* 30*nc, (iz,ic) iz-fastest */
    int nion;
    const int *ion_iz, *ion_ii;       /* nion, ascending (iz, ii) */
    const double *oc, *ginv;          /* nion*nc ion-fastest */
    int nline;
    const float *wl0, *gxs, *chilw;   /* nline, lambda-sorted */
    const int16_t *liz, *lii;
    const float *bf_ph1, *bf_ph2;     /* (6,7,30,30), (7,30,30) Fortran order */
    const double *ff_gff;             /* (81,41) Fortran order */
    int flags;                        /* bit0 no bb, bit1 no bf, bit2 no ff, bit3 no thomson */
} or_opac_in;

int or_opacity(const or_opac_in *in, double *cap64 /* ng*nc group-fastest */, float *cap32,
               double *sig, double *capgrey, double *capross, float *emiss);

/* ---- LTE EOS: gas_setup.f:101-160 (massfr2natomfr, no container fraction) +
 *      eos_update.f:23-45 + ionsmod.f:160-356 (ions_solve_eos, saha_nelec) ---- */
typedef struct {
    const int *ni;          /* 30: ionisation stages of element iz (ion_el(iz)%ni, incl. the bare nucleus) */
    const double *e;        /* 30*31, [iz-1][ii-1]: ionisation energy of stage ii [erg] */
    const double *q;        /* partition function, frozen at the host's first EOS call (H6) */
    const double *g1;       /* ground-level statistical weight glev(1) */
    const double *amass;    /* 30 element masses [amu] (elemdatamod.f) */
} or_eos_data;

/* Per massive cell: natom and natomfr from mass and mass fractions, then the Saha/secant
 * solve for nelec (in/out: the start value, H5) to relative accuracy acc (the reference's
 * 1e-8 when acc <= 0), then the ground-level populations oc/ginv of the listed ions.
 * massfr, natomfr: 30*nc iz-fastest; oc, ginv: nion*nc ion-fastest; niter: iterations used
 * (41 = not converged).  Returns 0, or 1 when the reference would stop (nelec <= 0 or an
 * invalid Saha normalisation). */
int or_eos(int nc, const or_eos_data *ed, const double *temp, const double *mass,
           const double *vol, const double *massfr, double acc, double *nelec, double *natom,
           double *natomfr, int nion, const int *ion_iz, const int *ion_ii, double *oc,
           double *ginv, int *niter);

/* ---- NLTE hydrogen: GAS/physical_opacity_nlte.f:3-708 ---- */
typedef struct {
    /* one NLTE species (H I), the permanent arrays of read_nlte_data.f:148-233; levels and
     * lines 1-based in the index arrays */
    int nlev, nlin, ncoll;
    const double *glev;                       /* nlev */
    const int *llw, *lup;                     /* nlin, lower level first */
    const double *wl0, *f, *gl, *gu;          /* nlin; wl0 [Angstrom] from the level energies */
    const double *aul, *bul, *blu;            /* nlin Einstein coefficients */
    const int *llw_coll, *lup_coll;           /* ncoll */
    const double *coll_c;                     /* 4*ncoll fit coefficients C0..C3 */
    const double *dele;                       /* ncoll transition energy [eV] */
    const int *lev_pi;                        /* nlev */
    const double *pi_c, *rr_c;                /* 4*nlev fits (PI, RR) */
    const double *dele_pi;                    /* nlev [eV] */
    /* per cell */
    const double *radtemp;                    /* nc */
    const double *ndens;                      /* 2*nc: number densities of H I and H II
                                                 (nlte_ndens(i,1:2,1), eos_update.f:52-57) */
} or_nlte_in;

/* physical_opacity_nlte for nlte_nelem = 1: per massive cell the rate matrix of the H I
 * levels (radiative rates with J at the group centre, electron-impact rates, photoionisation
 * and radiative recombination) solved for the level populations (LU with partial pivoting,
 * LAPACK dgetf2's unblocked algorithm, and dgetrs's substitutions; the reference inverts the
 * same matrix with dgetrf/dgetri and multiplies), the H lines' NLTE absorption and emission,
 * the other elements' lines in LTE, bf and ff, each added to cap, capemit and emiss as the
 * reference does; then the Planck means of cap and capemit (opacity_planckmean.f:18-24) and
 * the Rosseland mean.  The line list must hold no H I line (the reference's LTE list
 * excludes NLTE species).  npop (nlev*nc, level-fastest, may be NULL) receives the
 * normalised populations.  Returns 0, or 1 when a pivot is zero (the reference stops). */
int or_opacity_nlte(const or_opac_in *in, const or_nlte_in *nl, double *cap64, float *cap32,
                    float *capemit32, double *sig, double *capgrey, double *capemitgrey,
                    double *capross, float *emiss, double *npop);

/* the rate matrix (column-major n x n) and right-hand side of cell i, before the solve */
int or_nlte_rate_matrix(const or_opac_in *in, const or_nlte_in *nl, int i, double *a, double *b);

/* LU factorisation with partial pivoting (dgetf2) of the column-major n x n matrix a, and the
 * solution of a x = b in place (dgetrs, no transpose).  Returns 0 or the first zero pivot. */
int or_lu_solve(int n, double *a, int *ipiv, double *b);

/* ---- DDMC prep: leakage_opacity11.f90:30-171, emission_probability.f90:39-62 ---- */
void or_leakage(const or_model *m, double *opaclump);
void or_emission_probability(const or_model *m, double *emitprob);
int or_emitgroup(const or_model *m, double r, int ic);                       /* GRID/emitgroup.f:3-54 */

/* nltemod.f:231-260 nlte_rad_est: radtemp (in/out, nc) from the tally grd_tally (nc) and the
 * cell volumes; jrad (ng*nc group-fastest) at the group centres of wl (ng+1) */
void or_nlte_rad_est(int nc, int ng, const double *wl, const double *vol, const double *tally,
                     double *radtemp, double *jrad);

/* ---- sources: sourceenergy.f90:23-50, sourcenumbers.f90:26-169 (nmpi=1, no surface source) ---- */
void or_sourceenergy(int nc, const double *rcell, const double *radtemp, double lbol,
                     double *emit, double *sthermal);
int64_t or_sourcenumbers(int nc, const double *emit, int64_t nstot, int *nvol);

/* ---- one Monte Carlo iteration: interior_source + source_transformdirection +
 *      particle_advance (transport11/diffusion11) + fluxtally, Philox streams
 *      key (seed, iteration), counter (particle, block, phase 0=source 1=transport) ---- */
int or_iteration(const or_model *m, uint32_t seed, uint32_t iteration, or_result *res);

/* n draws of a particle's random stream: the sequential stream (step = 0; sources) or that
 * of transport step `step` >= 1 */
void or_stream_draws(uint32_t seed, uint32_t iter, uint32_t pid, uint32_t phase,
                     uint32_t step, int n, double *out);

/* source sampling only (interior_source.f90:57-168 + source_transformdirection.f90:23-46) */
int or_sources(const or_model *m, uint32_t seed, uint32_t iteration,
               double *x, double *mu, double *om, double *e, double *e0, double *wl);

#ifdef __cplusplus
}
#endif
#endif
