/*
 * lb3d_gpu.h -- C ABI of the MI355X (gfx950) LB3D timestep library
 * (lb3d_amd/liblb3d_gpu.so).
 *
 * This is the drop-in boundary for the per-timestep lattice update of LB3D.
 * The reference's interface is a set of argument-less Fortran module
 * procedures that mutate the global lattice whole_N:
 *   lb3d_time_loop()          code/lb3d_time_loop.F90:42-97
 *   lb3d_bc_init(stage)       code/lb3d_bc.F90:48-90
 *   lb3d_bc_apply(stage)      code/lb3d_bc.F90:95-155
 *   lb3d_advection()          code/lb3d_advection.F90:40-50
 *   lb3d_collision()          code/lb3d_collision.F90:41-47
 *   lb3d_force_apply(stage)   code/lb3d_force.F90:98-121
 *   lbe_sanity_check(N, ...)  code/lb3d_sanity.F90:79-384
 * Each entry point below replaces one of them; the Fortran host binds them
 * with ISO_C_BINDING (lb3d_amd/fortran/lb3d_gpu_module.F90) and keeps the
 * reference's names and stage order.  The lattice lives in device memory,
 * owned by the context; the host's array-of-structures whole_N is touched
 * only at upload/download.
 *
 * Conventions: every function returns 0 on success and a nonzero LB3D_E*
 * code on failure; lb3d_gpu_last_error() then describes the failure.  The
 * reference aborts through abend() -> MPI_Abort (code/lb3d_mpi.F90:670-676);
 * the Fortran host does the same on a nonzero code.  All work of a context
 * runs in order on one HIP stream owned by the context; upload, download and
 * diagnostics synchronise that stream.  One host thread per context.
 */
#ifndef LB3D_GPU_H
#define LB3D_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LB3D_ABI_VERSION 1

/* return codes */
#define LB3D_OK 0
#define LB3D_EINVAL 1   /* invalid argument or parameter combination */
#define LB3D_EHIP 2     /* a HIP runtime call failed */
#define LB3D_ESTATE 3   /* call out of the stage order (e.g. collision before advection) */
#define LB3D_ENOMEM 4   /* device or host allocation failed */

/* lb3d_params.precision */
#define LB3D_FP64 0     /* plain FP64 populations, the reference's arithmetic */
#define LB3D_FP32_DEV 1 /* FP32 deviations n_i - w_i*rho0 from the rest equilibrium */

/* lb3d_params.boundary */
#define LB3D_BC_PERIODIC 0     /* boundary_cond = 0 */
#define LB3D_BC_LEES_EDWARDS 1 /* boundary_cond != 0 and inv_fluid = 5 */

/* Physics and lattice parameters (bind(C) compatible, no padding holes).
 * Field meanings follow the reference's namelists (code/lb3d_config.F90:
 * 244-297, 319-335); omega = 1/tau is derived (code/lb3d_config.F90:886-889). */
typedef struct lb3d_params {
    int32_t nx, ny, nz;       /* global lattice (tnx, tny, tnz); one GPU holds it all */
    int32_t ncomp;            /* 2 = binary (-DNOSURFACTANT), 3 = ternary (default build) */
    int32_t bdist;            /* equilibrium form of code/lb3d_bdist.F90, 0-6: 0 Luo, 2 third-order
                                 Taylor (the fast forms), 1 Hudong, 3 second order, 4/5 Guo with
                                 the body-force term, 6 fourth-order Luo; batches take 0 and 2 */
    int32_t precision;        /* LB3D_FP64 or LB3D_FP32_DEV */
    int32_t boundary;         /* LB3D_BC_PERIODIC or LB3D_BC_LEES_EDWARDS (binary only) */
    int32_t restore;          /* 1: shear_sum was read from a checkpoint (code/lb3d_bc.F90:67-72) */
    double amass[3];          /* amass_r, amass_b, amass_s */
    double tau[3];            /* tau_r, tau_b, tau_s */
    double tau_d, beta, d_0;  /* dipole relaxation time, inverse temperature, dipole modulus */
    double g_br, g_bs, g_ss;  /* Shan-Chen couplings */
    double shear_u;           /* LE wall speed as in the input file (x tnx/(tnx-1) applied here) */
    double shear_omega;       /* LE oscillation frequency */
    double shear_sum;         /* LE offset accumulator: checktopo value on restore, else 0 */
    double g_accn[3];         /* body force acceleration (g_accn_x, g_accn_y, g_accn) */
    int32_t g_accn_min[3];    /* body-force box, global 1-based inclusive (x, y, z); */
    int32_t g_accn_max[3];    /*   max 0 means the lattice extent (code/lb3d_config.F90:931-933) */
    double rho0[3];           /* FP32_DEV reference densities (0 picks 0.7) */
    /* walls (rock sites come with the uploaded sites' rock_state; binary only):
     * wall couplings and relaxation weights (code/lb3d_config.F90:260-261) and
     * the colours rock sites carry (lbe_init_rock_colour, code/lb3d_ic_rock.F90:146-189) */
    double g_wr, g_wb, tau_wr, tau_wb;
    double rock_colour_r, rock_colour_b;
    /* invasion after the advection (lbe_invade, code/lb3d_invasion.F90:45-129,
     * with boundary_cond != 0): inv_fluid 0 none, 12 the Zou-He flux boundary
     * on the z planes (lbe_invade_zouandheflux, :2143-2457; periodic lattices
     * without walls); inv_type -1 (the flux is a mass flow), 0 or 1 (with the
     * corrected diagonals); inv_flux = (pr, pb, pg), the z flux of red, blue
     * and surfactant */
    int32_t inv_fluid, inv_type;
    double inv_flux[3];
} lb3d_params;

typedef struct lb3d_ctx lb3d_ctx;

/* Version of this ABI; the Fortran and Python bindings check it. */
int lb3d_gpu_abi_version(void);

/* Text of the last error on the calling thread ("" if none). */
const char *lb3d_gpu_last_error(void);

/* Create a context on the current HIP device: validates the parameters and
 * allocates the device lattice.  Replaces the allocation of whole_N at init
 * stage 3 (code/lb3d_lattice.F90:244-309). */
int lb3d_gpu_create(const lb3d_params *p, lb3d_ctx **out);
int lb3d_gpu_destroy(lb3d_ctx *ctx);

/* Copy the host lattice to the device.  whole_N is the reference's AoS array
 * of lbe_site records (code/lb3d_lattice.F90:35-83), Fortran order with x
 * fastest, extent [1-halo, n+halo] in each dimension; site_doubles is 49
 * (binary) or 79 (ternary).  Populations, (ternary) db and rock_state of the
 * interior are read; the halo, rho, u and da are ignored.  A site with
 * rock_state != 0 is a wall (code/lb3d_bc_bounce.F90:53-90): periodic binary
 * lattices only, which then run the two-pass plan.
 * Sets the step counter nt = nt0 (tstart). */
int lb3d_gpu_upload_sites(lb3d_ctx *ctx, const double *whole_N, int32_t halo_extent,
                          int32_t site_doubles, int32_t nt0);
/* Copy the lattice back into whole_N: interior populations, the moments of
 * the last advection (rho; u is the momentum of the current populations),
 * da and db; the halo is filled with periodic images of the interior (the
 * layout the reference's dumps see, code/lb3d_mpi.F90:788-858). */
int lb3d_gpu_download_sites(lb3d_ctx *ctx, double *whole_N, int32_t halo_extent,
                            int32_t site_doubles);

/* Seeded synthetic state on the device (not a reference routine): rest
 * equilibrium n_i = w_i*rho with rho = rho_mean[c] + amp[c]*(2u-1), u a
 * counter-based hash of (seed, c, site); ternary dipoles uniform on the unit
 * sphere when dipoles != 0.  Sets nt = 0. */
int lb3d_gpu_init_noise(lb3d_ctx *ctx, uint64_t seed, const double rho_mean[3],
                        const double amp[3], int32_t dipoles);

/* Stage-4 boundary initialisation (code/lb3d_bc.F90:48-90): for LE rescales
 * shear_u by tnx/(tnx-1), applies the restore correction and the first
 * le_neighbours.  Call once after upload/init_noise. */
int lb3d_gpu_bc_init(lb3d_ctx *ctx);

/* The reference's per-step call surface, in the order of
 * code/lb3d_time_loop.F90:58-92:
 *   force_apply(6) bc_apply(6, nt) advection() bc_apply(8, nt) force_apply(8) collision()
 * Out-of-order calls return LB3D_ESTATE.  Periodic boundaries are index wraps
 * inside the kernels, so bc_apply only does work for Lees-Edwards. */
int lb3d_gpu_force_apply(lb3d_ctx *ctx, int32_t stage);          /* code/lb3d_force.F90:98-121 (no-op) */
int lb3d_gpu_bc_apply(lb3d_ctx *ctx, int32_t stage, int32_t nt); /* code/lb3d_bc.F90:95-155 */
int lb3d_gpu_advection(lb3d_ctx *ctx);                           /* code/lb3d_advection.F90:40-151 */
int lb3d_gpu_collision(lb3d_ctx *ctx);                           /* code/lb3d_collision.F90:41-291 */

/* nsteps whole timesteps nt+1 .. nt+nsteps (the body of lb3d_time_loop,
 * code/lb3d_time_loop.F90:58-92), replayed from a captured HIP graph. */
int lb3d_gpu_run_steps(lb3d_ctx *ctx, int32_t nsteps);
/* Timing support (not a reference routine): lb3d_gpu_run_steps between two
 * HIP events recorded on the context's stream; *ms is the elapsed time
 * between them.  Synchronises. */
int lb3d_gpu_run_steps_timed(lb3d_ctx *ctx, int32_t nsteps, double *ms);

/* Autotuner (not a reference routine): times every execution plan of this
 * lattice -- the two-pass kernels and, for binary lattices, the one-pass
 * z-march kernel over its tile shapes and z-chunks, then the XCD chunking of
 * the best -- on trial_steps trial steps each, keeps the fastest and writes a
 * text report (may be NULL).  The lattice state is unchanged. */
int lb3d_gpu_autotune(lb3d_ctx *ctx, int32_t trial_steps, char *report, int32_t report_len);
/* Choose the execution plan directly: fused = 0 (two-pass), 1 (one-pass
 * kernel with tile tx x ty and z-chunk lz, look-ahead plane staged in LDS),
 * 2 (the same with the look-ahead plane in registers) or 3 (no look-ahead
 * copy: the collide pulls its plane again); 1-3 binary with bdist 2 only,
 * 3 also ternary (periodic, bdist 0 or 2).
 * xcd_group: tiles that one XCD (and its L2) takes in a row; 1 is the plain
 * round-robin order, 0 one contiguous run of tiles per XCD.  Results do not
 * depend on the plan. */
int lb3d_gpu_set_plan(lb3d_ctx *ctx, int32_t fused, int32_t tx, int32_t ty, int32_t lz,
                      int32_t xcd_group);

/* Current step counter nt. */
int lb3d_gpu_get_nt(lb3d_ctx *ctx, int32_t *nt);

/* Sanity reductions in FP64 (code/lb3d_sanity.F90:79-224): mass[c] = sum of
 * rho_c, momentum = total mass flow, counts = {sites with a NaN population,
 * sites with a negative population, sites with |u_c| > 0.1}.  Synchronises. */
int lb3d_gpu_diagnostics(lb3d_ctx *ctx, double mass[3], double momentum[3], int64_t counts[3]);

/* The full sanity check of code/lb3d_sanity.F90:79-384 (one rank), in FP64
 * on the current populations.  Per species c (0 r, 1 b, 2 s) and axis a:
 * the velocity v = amass_c * sum_i g_i n_c,i c_i,a / rho_c of every site with
 * |rho_c| > 1e-12 is summed (vel_sum; divide by the site count for the
 * reference's average), and its largest magnitude is kept with the global
 * 1-based (x, y, z) of that site (max_pos[c][a]; among equal maxima the site
 * last in the reference's loop, x outermost and z innermost; 0 if none). */
typedef struct lb3d_sanity {
    double mass[3];             /* total density per component */
    double momentum[3];         /* total mass flow, sum_c amass_c * P_c */
    double vel_sum[3][3];
    double max_vel[3][3];
    int32_t max_pos[3][3][3];
    int32_t pad0;
    int64_t fluid_sites;
    int64_t nan_sites;          /* sites with a NaN population (found_NaN if > 0) */
    int64_t negative_sites;     /* sites with a negative population */
    int64_t fast_sites;         /* sites where some species is faster than 0.1 */
} lb3d_sanity;
int lb3d_gpu_sanity(lb3d_ctx *ctx, lb3d_sanity *out);

/* Lees-Edwards scalars after the last le_neighbours (code/lb3d_bc_leesedwards.F90:86-140). */
int lb3d_gpu_le_state(lb3d_ctx *ctx, double *shear_sum, int32_t *lelo, double *lefrac,
                      double *shear_u);

/* Populations of the interior as the solver holds them, in FP64:
 * out[((c*19 + i)*nz + z)*ny*nx + y*nx + x] (FP32_DEV: w_i*rho0 + deviation). */
int lb3d_gpu_get_populations(lb3d_ctx *ctx, double *out);

/* Densities rho_c of the interior in FP64, out[(c*nz + z)*ny*nx + y*nx + x]
 * (the od/wd/sur fields of code/lb3d_io_dump_data.F90:922-1166 without
 * amass).  Synchronises. */
int lb3d_gpu_get_density(lb3d_ctx *ctx, double *out);

/* Momenta P_c = sum_i g_i n_c,i c_i of the interior in FP64 (the u field of
 * lbe_site, lb3d_update_velocity, code/lb3d_analysis.F90:78-102),
 * out[((c*3 + a)*nz + z)*ny*nx + y*nx + x].  Synchronises. */
int lb3d_gpu_get_momentum(lb3d_ctx *ctx, double *out);

/* Lees-Edwards x-halo planes that the pre-advection boundary pass builds from
 * the current populations with the current offset (lelo, lefrac of
 * lb3d_gpu_le_state): the interpolated, velocity-shifted and clamped planes of
 * le_halo_exchange (code/lb3d_bc_leesedwards.F90:659-1004), in FP64,
 * out[((f*2 + c)*19 + i)*nz*ny + z*ny + y], f = 0 the x = 0 plane, f = 1 the
 * x = nx+1 plane (y, z 0-based interior).  Binary Lees-Edwards lattices only;
 * needs a completed timestep.  Synchronises. */
int lb3d_gpu_le_halo(lb3d_ctx *ctx, double *out);

/* The arr observable of the reference's dumps (dump_arrows,
 * code/lb3d_io_dump_data.F90:1308-1460) in FP64: per site
 * (sum_c amass_c (P_c - F_c rho_c / 2)) / max(10.e-9, sum_c amass_c rho_c),
 * less g_accn/2 inside the body-force box, with the Shan-Chen forces F_c of
 * the state (from the densities of the last advection and the dipole field,
 * periodic in x as the reference's dumps see it) and rho_c, P_c summed from
 * the populations; out[3*(x + nx*(y + ny*z)) + a] (the order of the dump).
 * Needs a completed timestep (LB3D_ESTATE otherwise).  Synchronises. */
int lb3d_gpu_get_arrows(lb3d_ctx *ctx, double *out);

/* Wait for all work queued on the context's stream. */
int lb3d_gpu_synchronize(lb3d_ctx *ctx);

/* Timing support: runs nsteps timesteps with HIP events around every kernel
 * on the context's stream and returns the mean duration per launch.
 * names/ms have room for max_kernels entries; *nkernels is set to the count. */
int lb3d_gpu_profile_steps(lb3d_ctx *ctx, int32_t nsteps, int32_t max_kernels,
                           const char **names, double *ms, int32_t *nkernels,
                           double *ms_per_step);

/* ---------------------------------------------------------------------------
 * One timestep that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  A
 * periodic lattice of parameters p advances one step of the two-pass plan
 * from f_in to f_out (the populations in the device layout of p->precision:
 * f[(c*19 + i)*ns + s], FP64 values or FP32 deviations n_i - w_i*rho0_c;
 * s = x + nx*(y + ny*z)); ternary lattices also read the dipole field db_in
 * and write the relaxed field to db_out (db[k*ns + s]).  All four are
 * caller-owned device buffers; the moments field lives in a static device
 * array of LB3D_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends,
 * and lattices whose moments need more return LB3D_EINVAL.  The step is
 * enqueued on stream (a hipStream_t; NULL = the default stream) and returns
 * without waiting.  Every call on one GPU shares that GPU's array: calls that
 * overlap in time on one GPU (other streams, other threads) race. */
#define LB3D_NOALLOC_SCRATCH_BYTES (64ull << 20)
int lb3d_gpu_step_noalloc(const lb3d_params *p, const void *f_in, void *f_out, const void *db_in,
                          void *db_out, void *stream);

/* ---------------------------------------------------------------------------
 * Batched ensemble (not a reference routine: the reference runs one lattice
 * per job).  Many independent periodic lattices of uneven sizes -- parameter
 * sweeps, seeds -- advanced together: each timestep is one persistent launch
 * per phase over all lattices, replayed from a captured HIP graph.  All
 * lattices share ncomp and precision; every other parameter is per lattice
 * (bdist 0-6, body force; no walls or invasion).
 */
typedef struct lb3d_batch lb3d_batch;

int lb3d_gpu_batch_create(int32_t nlat, const lb3d_params *params, lb3d_batch **out);
int lb3d_gpu_batch_destroy(lb3d_batch *b);
/* State of lattice k: seeded noise (as lb3d_gpu_init_noise) or the
 * reference's AoS lattice (as lb3d_gpu_upload_sites).  Both reset nt to 0. */
int lb3d_gpu_batch_init_noise(lb3d_batch *b, int32_t k, uint64_t seed, const double rho_mean[3],
                              const double amp[3], int32_t dipoles);
int lb3d_gpu_batch_upload_sites(lb3d_batch *b, int32_t k, const double *whole_N,
                                int32_t halo_extent, int32_t site_doubles);
/* nsteps timesteps of every lattice. */
int lb3d_gpu_batch_run_steps(lb3d_batch *b, int32_t nsteps);
/* lb3d_gpu_batch_run_steps between HIP events on the batch's stream; *ms is
 * the elapsed time (as lb3d_gpu_run_steps_timed). */
int lb3d_gpu_batch_run_steps_timed(lb3d_batch *b, int32_t nsteps, double *ms);
int lb3d_gpu_batch_get_nt(lb3d_batch *b, int32_t *nt);
int lb3d_gpu_batch_synchronize(lb3d_batch *b);
/* The full sanity check of lattice k (as lb3d_gpu_sanity). */
int lb3d_gpu_batch_sanity(lb3d_batch *b, int32_t k, lb3d_sanity *out);
/* Populations of lattice k in FP64 (layout as lb3d_gpu_get_populations). */
int lb3d_gpu_batch_get_populations(lb3d_batch *b, int32_t k, double *out);
/* Sanity reductions of lattice k (as lb3d_gpu_diagnostics). */
int lb3d_gpu_batch_diagnostics(lb3d_batch *b, int32_t k, double mass[3], double momentum[3],
                               int64_t counts[3]);
/* Plan: persistent launches with nwg workgroups (sequential = 0; nwg is capped
 * at the tile count and rounded up to a multiple of 8, one per XCD) taking
 * the work list in chunks of xcd_group items per XCD (as lb3d_gpu_set_plan),
 * or one lattice after another with the single-lattice kernels
 * (sequential = 1). */
int lb3d_gpu_batch_set_plan(lb3d_batch *b, int32_t sequential, int32_t nwg, int32_t xcd_group);
/* Collide-moments plan (FP32 ternary batches): each timestep is the dipole
 * relaxation, one collision kernel that also gathers the moments of the next
 * step's streamed state from the populations it has just written, and a
 * pass over the tile edges whose moments it cannot gather; the populations
 * are read once per step.  Tiles of about tile_cols columns and as many
 * rows as work-groups of tile_threads (256 or 512) hold (tile_cols = 0 turns
 * the plan off, back to set_plan's), z-chunks of about lz planes,
 * xcd_group work items per XCD chunk.  Results equal the other plans'
 * bitwise (the FP32 ternary moments are summed in one order by every plan).
 * Not a reference routine: the reference runs one lattice per job. */
int lb3d_gpu_batch_set_collide_moments(lb3d_batch *b, int32_t tile_cols, int32_t tile_threads, int32_t lz,
                                       int32_t xcd_group);
/* Groups of the persistent plan: consecutive lattices whose populations add
 * up to about group_bytes (a larger lattice alone) run their moments, dipole
 * and collide phases before the next group starts, so that the collide finds
 * the group's populations and moments in the memory-side cache (256 MiB on
 * MI355X).  0 = one group: each phase over the whole batch.  The results do
 * not depend on the grouping.  No counterpart in the reference (an execution
 * plan of the batched ensemble, which the reference does not have). */
int lb3d_gpu_batch_set_groups(lb3d_batch *b, uint64_t group_bytes);
/* Autotuner: times the sequential plan, persistent grid sizes, XCD chunks
 * and lattice groups on the batch's own state (unchanged) and keeps the
 * fastest. */
int lb3d_gpu_batch_autotune(lb3d_batch *b, int32_t trial_steps, char *report, int32_t report_len);
/* Mean HIP-event duration of the moments, dipole and collide phases (the
 * collide-moments plan: its edge moments, dipole and collide launches). */
int lb3d_gpu_batch_profile_steps(lb3d_batch *b, int32_t nsteps, double ms_phase[3],
                                 double *ms_per_step);

#ifdef __cplusplus
}
#endif
#endif
