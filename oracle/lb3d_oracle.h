/*
 * lb3d_oracle.h -- CPU restatement of the LB3D per-timestep lattice update.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP path is compared
 * against; it is never linked into, loaded by, or called from the product
 * library (lb3d_amd/liblb3d_gpu.so).  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.
 *
 * The state is the reference's own array-of-structures lattice: whole_N with a
 * one-site halo, Fortran order (x fastest), one `lbe_site` record per site
 * (code/lb3d_lattice.F90:35-83).  Offsets in doubles:
 *   component c (0=r,1=b,2=s) at 24*c: n[19], rho, u[3] (momentum), rock_colour
 *   ternary: da[3] at 72, db[3] at 75, rock_state at 78   (79 doubles)
 *   binary:  rock_state at 48                             (49 doubles)
 * All arithmetic is FP64 in the reference's operation order; build with
 * -ffp-contract=off so that no multiply-add is fused.
 */
#ifndef LB3D_ORACLE_H
#define LB3D_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_params {
    int32_t nx, ny, nz;      /* lattice size (one rank holds everything) */
    int32_t ncomp;           /* 2 = binary (-DNOSURFACTANT), 3 = ternary */
    int32_t bdist;           /* 0 (Luo) or 2 (third-order Taylor) */
    int32_t lees_edwards;    /* 0: periodic (boundary_cond=0); 1: LE (inv_fluid=5) */
    int32_t restore;         /* 1: shear_sum comes from a checkpoint */
    int32_t pad0;
    double amass[3], tau[3];
    double tau_d, beta, d_0;
    double g_br, g_bs, g_ss;
    double shear_u;          /* input value; bc_init applies tnx/(tnx-1) */
    double shear_omega;
    double shear_sum;        /* checktopo value on restore, else 0 */
    double g_accn[3];        /* g_accn_x, g_accn_y, g_accn */
    int32_t g_accn_min[3];   /* body-force box, global 1-based inclusive (x,y,z) */
    int32_t g_accn_max[3];
    /* walls (code/lb3d_config.F90:181-183, 260-261): wall couplings and
     * relaxation weights, rock colours of lbe_init_rock_colour */
    double g_wr, g_wb, tau_wr, tau_wb;
    double rock_colour_r, rock_colour_b;
    /* invasion at stage 8 (code/lb3d_invasion.F90:45-129): inv_fluid 12 is the
     * Zou-He flux boundary on the z planes (lbe_invade_zouandheflux), any other
     * value none; inv_type -1, 0 or 1; inv_flux = (pr, pb, pg), the z flux of
     * red, blue and surfactant */
    int32_t inv_fluid, inv_type;
    double inv_flux[3];
} or_params;

typedef struct or_state or_state;

/* Worker threads of the site loops (OpenMP build); results are identical for
 * every thread count because each site is computed on its own. */
void or_set_threads(int n);
int  or_threads(void);

or_state *or_create(const or_params *p);
void      or_destroy(or_state *o);
double   *or_sites(or_state *o);             /* whole_N, (nx+2)(ny+2)(nz+2)*site_doubles */
int       or_site_doubles(const or_state *o);
int       or_nt(const or_state *o);
void      or_set_nt(or_state *o, int nt);

/* stage-4 boundary init (code/lb3d_bc.F90:48-90) after the state is loaded */
void or_bc_init(or_state *o);
/* the reference's call surface, one stage at a time */
void or_bc_apply(or_state *o, int stage);    /* stages 6 and 8 */
/* the Zou-He flux rule on the 19 populations of one component at one site of
 * the z = 1 plane (low = 1) or the z = nz plane (low = 0) */
void or_zou_he_site(double *n, double uz, int itype, int low);
void or_advection(or_state *o);              /* stage 7 */
void or_collision(or_state *o);              /* stage 9 */
/* one timestep nt -> nt+1 (code/lb3d_time_loop.F90:58-92) and n of them */
void or_time_step(or_state *o);
void or_run(or_state *o, int nsteps);

/* obstacles (lbe_init_rock, code/lb3d_ic_rock.F90:50-77 with read_rock_xdrf_par,
 * code/lb3d_io_xdrf.F90:400-470, and lbe_init_rock_colour, :146-189): a site is
 * rock where obs (interior, x fastest) exceeds threshold, fluid elsewhere; rock
 * sites lose their populations (and ternary da, db) and take the rock colours
 * that are positive.  Call before or_bc_init. */
void or_init_rock(or_state *o, const double *obs, double threshold);

/* the pre-advection halo of the current state, offset not advanced */
void or_halo_refresh(or_state *o);
/* momentum transfer through x = minx and x = maxx (code/lb3d_io_stress.F90:38-173),
 * halved as dump_stress does; mean shear stress = (out[1] - out[0])/(ny*nz) */
void or_stress(const or_state *o, int minx, int maxx, double out[2]);

/* full sanity check (code/lb3d_sanity.F90:79-384); species c, axis a */
typedef struct or_sanity_report {
    double mass[3], momentum[3];
    double vel_sum[3][3];      /* sum of site velocities amass*P/rho where |rho| > 1e-12 */
    double max_vel[3][3];      /* largest |velocity component| */
    int32_t max_pos[3][3][3];  /* 1-based (x,y,z) of the last >= update; 0 if none */
    int32_t pad0;
    int64_t fluid_sites, nan_sites, negative_sites, fast_sites;
} or_sanity_report;
void or_sanity_full(const or_state *o, or_sanity_report *r);

/* LE scalars after the last le_neighbours call */
void or_le_state(const or_state *o, double *shear_sum, int *lelo, double *lefrac,
                 double *shear_u);

/* sanity-check reductions (code/lb3d_sanity.F90:79-224):
 * mass[c] = sum rho_c, mom[3] = total mass flow, counts = {NaN sites,
 * sites with a negative population, sites faster than 0.1} */
void or_sanity(const or_state *o, double mass[3], double mom[3], int64_t counts[3]);

/* observables of the reference dumps at stage 10, interior sites, x fastest
 * (code/lb3d_io_dump_data.F90): od, wd, sur are scalars (nx*ny*nz),
 * dir and arr are 3-vectors interleaved per site (3*nx*ny*nz), vel scalar. */
void or_dump_od(const or_state *o, double *out);
void or_dump_wd(const or_state *o, double *out);
void or_dump_sur(const or_state *o, double *out);
void or_dump_dir(const or_state *o, double *out);
void or_dump_vel(const or_state *o, double *out);
void or_dump_arr(const or_state *o, double *out);

/* equilibrium distribution (code/lb3d_bdist.F90:175-389), bdist 0-6; the
 * _dd form takes Guo's forcing term dd (bdist 4 and 5; 0 for or_boltz_dist) */
void or_boltz_dist(int bdist, double ux, double uy, double uz, double dux,
                   double duy, double duz, double *boltz19);
void or_boltz_dist_dd(int bdist, double ux, double uy, double uz, double dux, double duy, double duz,
                      double ddx, double ddy, double ddz, double *boltz19);
/* Shan-Chen force at one interior site from the current state */
void or_sc_force(const or_state *o, int x, int y, int z, double f[3][3]);

/* Seeded synthetic initial state shared by the oracle and the GPU library:
 * rest-equilibrium populations n_i = W_i * rho with
 * rho = rho_mean[c] + amp[c]*(2u-1), u from lb3d_hash_uniform(seed, c, site);
 * ternary dipoles uniform on the unit sphere from streams 3 and 4.
 * Writes the interior of whole_N; halos are filled by or_bc_init. */
void or_init_noise(or_state *o, uint64_t seed, const double rho_mean[3],
                   const double amp[3], int dipoles);
double or_hash_uniform(uint64_t seed, uint32_t stream, uint64_t site);

#ifdef __cplusplus
}
#endif
#endif
