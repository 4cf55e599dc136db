/*
 * lb3d_oracle.c -- FP64 CPU restatement of LB3D's per-timestep update.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product; see the header).
 *
 * Each routine follows the reference routine named in its comment, in the
 * reference's loop and operation order, on the reference's own AoS lattice
 * (whole_N with a one-site halo).  Files cited are under the reference's
 * code/ directory.  Build with -O2 -ffp-contract=off.
 *
 * Pinning (what checks each function; the tests live in tests/):
 *   boltz_dist (bdist 0, 2)  : moment closed forms (sum g*feq = 1,
 *                              sum g*c*feq = u+du, rest/axis values at u=0)
 *                              and the reference's golden dumps
 *   advect + moments         : streaming is a permutation (brute force on a
 *                              tiny lattice), exact mass, golden dumps
 *   sc forces (binary)       : uniform rho => F=0; linear psi field =>
 *                              F = -12 g_br psi(x) grad psi; golden arr dump
 *   sc forces (ternary), dipole relaxation : gyroid golden dumps (od, wd,
 *                              sur, dir, vel, arr after 10 steps), and the
 *                              h=0 closed form of the dipole update
 *   collision (BGK)          : rest state fixed point, per-site mass exact,
 *                              golden dumps
 *   LE halo                  : leesEdwards golden dumps (od, wd, vel, arr
 *                              after 10 steps), uniform-fluid halo equals
 *                              feq(rho, (0,0,-+2U)), z-linear field shift
 *   body force               : bodyForce golden dumps
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include "lb3d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* D3Q19 vectors and degeneracies, code/lb3d_lattice.F90:137-164 (0-based) */
static const int cx[19] = {1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
static const int cy[19] = {0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0};
static const int cz[19] = {0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, 0};
static const int gdeg[19] = {2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
/* opposite vector, code/lb3d_lattice.F90:173-177 (0-based) */
static const int bounce_of[19] = {1, 0, 3, 2, 5, 4, 11, 10, 13, 12, 7, 6, 9, 8, 17, 16, 15, 14, 18};
/* float literals widened to double, e.g. dble(10.e-9) in code/lb3d_collision.F90:586 */
#define CLAMP_RHO 9.99999993922529e-09
#define CLAMP_SURF 9.999999747378752e-05

struct or_state {
    or_params p;
    int sd;            /* doubles per site */
    int lx, ly, lz;    /* nx+2 ... */
    double omega[3], omega_d;
    double shear_u, shear_tmp, shear_sum, lefrac;
    int lelo;
    int nt;
    double *N;         /* whole_N */
    double *buf;       /* advection source copy */
    double *force;     /* f(nx,ny,nz,3) per component */
};

#define OFF_N(c) (24 * (c))
#define OFF_RHO(c) (24 * (c) + 19)
#define OFF_U(c) (24 * (c) + 20)
#define OFF_DA 72
#define OFF_DB 75
#define OFF_RC(c) (24 * (c) + 23) /* rock_colour_r / rock_colour_b */

static inline int off_rock(const or_state *o) { return o->p.ncomp == 3 ? 78 : 48; }
static inline size_t sidx(const or_state *o, int x, int y, int z)
{
    return ((size_t)x + (size_t)o->lx * ((size_t)y + (size_t)o->ly * (size_t)z)) * o->sd;
}
static inline double *site(or_state *o, int x, int y, int z) { return o->N + sidx(o, x, y, z); }
static inline const double *csite(const or_state *o, int x, int y, int z)
{
    return o->N + sidx(o, x, y, z);
}

or_state *or_create(const or_params *p)
{
    or_state *o = (or_state *)calloc(1, sizeof(or_state));
    if (!o) return NULL;
    o->p = *p;
    o->sd = p->ncomp == 3 ? 79 : 49;
    o->lx = p->nx + 2;
    o->ly = p->ny + 2;
    o->lz = p->nz + 2;
    size_t n = (size_t)o->lx * o->ly * o->lz * o->sd;
    o->N = (double *)calloc(n, sizeof(double));
    o->buf = (double *)calloc(n, sizeof(double));
    o->force = (double *)calloc((size_t)p->nx * p->ny * p->nz * 3 * 3, sizeof(double));
    if (!o->N || !o->buf || !o->force) {
        or_destroy(o);
        return NULL;
    }
    /* omega = 1/tau, code/lb3d_config.F90:886-889 */
    for (int c = 0; c < 3; c++) o->omega[c] = 1.0 / p->tau[c];
    o->omega_d = 1.0 / p->tau_d;
    o->shear_u = p->shear_u;
    o->shear_tmp = p->shear_u;
    o->shear_sum = p->shear_sum;
    return o;
}

void or_destroy(or_state *o)
{
    if (!o) return;
    free(o->N);
    free(o->buf);
    free(o->force);
    free(o);
}

/* Worker threads of the site loops (OpenMP; every site is computed exactly
 * as with one thread, so results do not depend on the count). */
void or_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int or_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

double *or_sites(or_state *o) { return o->N; }
int or_site_doubles(const or_state *o) { return o->sd; }
int or_nt(const or_state *o) { return o->nt; }
void or_set_nt(or_state *o, int nt) { o->nt = nt; }

void or_le_state(const or_state *o, double *shear_sum, int *lelo, double *lefrac, double *shear_u)
{
    *shear_sum = o->shear_sum;
    *lelo = o->lelo;
    *lefrac = o->lefrac;
    *shear_u = o->shear_u;
}

/* ---------------------------------------------------------------------------
 * Equilibrium distribution, code/lb3d_bdist.F90:175-370 (COMPAT_BDIST build).
 * bdist 0 = Luo operator (:214-234), bdist 2 = third-order Taylor (:301-370).
 */
/* The other equilibrium forms of code/lb3d_bdist.F90 (COMPAT_BDIST):
 * 3 second order in (u, du) (:372-391), 4 and 5 Guo's second and fourth order
 * with the forcing term dd (:392-455), 6 the fourth-order Luo form (:464-482),
 * 1 the iterated exponential (Hudong) form (:237-300); 7 computes nothing in
 * the reference and is rejected.  Integer powers are
 * products, x**4 as (x*x)*(x*x); divisions by T are divisions, as written. */
static void boltz_general(int bdist, double ux, double uy, double uz, double dux, double duy, double duz,
                          double ddx, double ddy, double ddz, double *b)
{
    const double T = 1.0 / 3.0, T2 = T * T, T3 = T2 * T;
    const double w = 1.0 / 36.0, w0 = 1.0 / 3.0;
    if (bdist == 3) {
        double uke = 0.5 * (ux * ux + uy * uy + uz * uz);
        double udotdu = ux * dux + uy * duy + uz * duz;
        for (int s = 0; s < 18; s++) {
            double cu = cx[s] * ux + cy[s] * uy + cz[s] * uz;
            double cdu = cx[s] * dux + cy[s] * duy + cz[s] * duz;
            double q = cu / T;
            b[s] = w * (1.0 + q + 0.5 * (q * q) - uke / T + cdu / T + (cu * cdu) / T2 - udotdu / T);
        }
        b[18] = w0 * (1.0 - uke / T - udotdu / T);
        return;
    }
    if (bdist == 4 || bdist == 5 || bdist == 6) {
        double tux = ux + dux, tuy = uy + duy, tuz = uz + duz;
        double uke = 0.5 * (tux * tux + tuy * tuy + tuz * tuz);
        double dke = bdist == 6 ? 0.0 : 0.5 * (ddx * ddx + ddy * ddy + ddz * ddz);
        double t1x = bdist == 5 ? tux - ddx : tux, t1y = bdist == 5 ? tuy - ddy : tuy,
               t1z = bdist == 5 ? tuz - ddz : tuz;
        double u1ke = 0.5 * (t1x * t1x + t1y * t1y + t1z * t1z);
        for (int s = 0; s < 18; s++) {
            double cu = cx[s] * tux + cy[s] * tuy + cz[s] * tuz;
            double q = cu / T;
            if (bdist == 6) {
                double q2 = q * q, h = uke / T;
                b[s] = w * (1.0 + q + 0.5 * q2 - uke / T - (uke * cu) / T2 + (q2 * q) / 6.0 + 0.5 * (h * h) -
                            (0.5 * ((cu * cu) * uke)) / T3 + (q2 * q2) / 24.0);
                continue;
            }
            double cd = cx[s] * ddx + cy[s] * ddy + cz[s] * ddz;
            double r = cd / T;
            double v = 1.0 + q + 0.5 * (q * q - r * r) - (uke - dke) / T;
            if (bdist == 5) {
                double c1 = cx[s] * t1x + cy[s] * t1y + cz[s] * t1z;
                double q1 = c1 / T, q12 = q1 * q1, h1 = u1ke / T;
                v = v - (uke * c1) / T2 + (q12 * q1) / 6.0 + 0.5 * (h1 * h1) - (0.5 * ((c1 * c1) * u1ke)) / T3 +
                    (q12 * q12) / 24.0;
            }
            b[s] = w * v;
        }
        double h = uke / T;
        if (bdist == 4)
            b[18] = w0 * (1.0 - (uke - dke) / T);
        else if (bdist == 5)
            b[18] = w0 * (1.0 - (uke - dke) / T + 0.5 * (h * h));
        else
            b[18] = w0 * (1.0 - uke / T + 0.5 * (h * h));
        return;
    }
    /* bdist == 1 */
    double uxf = ux + dux, uyf = uy + duy, uzf = uz + duz;
    double uke_eq = 0.5 * (uxf * uxf + uyf * uyf + uzf * uzf);
    double u_mag = sqrt(2.0 * uke_eq);
    int ifactor = (int)(200.0 * u_mag);
    double n_iter = (double)(ifactor > 2 ? ifactor : 2);
    double delta = sqrt(1.0 / n_iter);
    double uxe = uxf, uye = uyf, uze = uzf, rho_eq = 1.0;
    for (int it = 1; it <= (int)n_iter; it++) {
        double ex[3] = {exp(uxe / T), exp(uye / T), exp(uze / T)};
        double expek = 1.0 / exp(uke_eq / T);
        for (int s = 0; s < 18; s++) {
            const int cv[3] = {cx[s], cy[s], cz[s]};
            double a = 1.0;
            for (int k = 0; k < 3; k++) {
                double pk = cv[k] == 1 ? ex[k] : (cv[k] == -1 ? 1.0 / ex[k] : 1.0);
                a = k == 0 ? pk : a * pk;
            }
            b[s] = ((rho_eq * w) * a) * expek;
        }
        b[18] = (rho_eq * w0) * expek;
        double r1 = 0.0, px = 0.0, py = 0.0, pz = 0.0;
        for (int s = 0; s < 18; s++) {
            r1 = r1 + b[s];
            px = px + b[s] * cx[s];
            py = py + b[s] * cy[s];
            pz = pz + b[s] * cz[s];
        }
        double rho_tt = r1 + b[18];
        rho_eq = rho_eq / (1.0 + (delta * (rho_tt - 1.0)) / rho_eq);
        uxe = uxe + delta * (uxf - px);
        uye = uye + delta * (uyf - py);
        uze = uze + delta * (uzf - pz);
        uke_eq = 0.5 * (uxe * uxe + uye * uye + uze * uze);
    }
}

void or_boltz_dist_dd(int bdist, double ux, double uy, double uz, double dux, double duy, double duz,
                      double ddx, double ddy, double ddz, double *boltz)
{
    if (bdist == 0 || bdist == 2)
        or_boltz_dist(bdist, ux, uy, uz, dux, duy, duz, boltz);
    else
        boltz_general(bdist, ux, uy, uz, dux, duy, duz, ddx, ddy, ddz, boltz);
}

void or_boltz_dist(int bdist, double ux, double uy, double uz, double dux, double duy,
                   double duz, double *boltz)
{
    if (bdist != 0 && bdist != 2) {
        boltz_general(bdist, ux, uy, uz, dux, duy, duz, 0.0, 0.0, 0.0, boltz);
        return;
    }
    const double T = 1.0 / 3.0;
    const double holdit = 1.0 / T;
    const double w_move = 1.0 / 36.0, w_rest = 1.0 / 3.0;
    if (bdist == 0) {
        double uke = 0.5 * (ux * ux + uy * uy + uz * uz);
        double udotdu = ux * dux + uy * duy + uz * duz;
        double uholdit = uke * holdit;
        for (int s = 0; s < 18; s++) {
            double cdotu = cx[s] * ux + cy[s] * uy + cz[s] * uz;
            double cdotdu = cx[s] * dux + cy[s] * duy + cz[s] * duz;
            double choldit = cdotu * holdit;
            double cdholdit = cdotdu * holdit;
            double ch2 = choldit * choldit;
            boltz[s] = w_move * (1.0 + choldit + 0.5 * ch2 - uholdit + (choldit * ch2) / 6.0 -
                                 uholdit * choldit + cdholdit + choldit * cdholdit -
                                 udotdu * holdit + 0.5 * ch2 * cdholdit -
                                 choldit * udotdu * holdit - uholdit * cdholdit);
        }
        boltz[18] = w_rest * (1.0 - uke / T - udotdu / T);
        return;
    }
    /* bdist == 2 */
    double tux = ux + dux, tuy = uy + duy, tuz = uz + duz;
    double tuxt = tux * holdit, tuyt = tuy * holdit, tuzt = tuz * holdit;
    double uke = 0.5 * (tux * tuxt + tuy * tuyt + tuz * tuzt);
    double cst = 1.0 - uke;
    double const1 = w_move * cst, const2 = w_move * cst;
    double w1by2 = w_move / 2.0, w2by2 = w_move / 2.0;
    double w1by6 = w_move / 6.0, w2by6 = w_move / 6.0;
    double cu, even, odd;
#define PAIR(CU, C, WBY2, WBY6, IP, IM)               \
    cu = (CU);                                        \
    even = C + WBY2 * (cu * cu);                      \
    odd = C * cu + WBY6 * (cu * cu * cu);             \
    boltz[IP] = even + odd;                           \
    boltz[IM] = even - odd;
    PAIR(tuxt, const1, w1by2, w1by6, 0, 1)
    PAIR(tuyt, const1, w1by2, w1by6, 2, 3)
    PAIR(tuzt, const1, w1by2, w1by6, 4, 5)
    PAIR(tuxt + tuyt, const2, w2by2, w2by6, 6, 11)
    PAIR(tuxt - tuyt, const2, w2by2, w2by6, 7, 10)
    PAIR(tuxt + tuzt, const2, w2by2, w2by6, 8, 13)
    PAIR(tuxt - tuzt, const2, w2by2, w2by6, 9, 12)
    PAIR(tuyt + tuzt, const2, w2by2, w2by6, 14, 17)
    PAIR(tuyt - tuzt, const2, w2by2, w2by6, 15, 16)
#undef PAIR
    boltz[18] = w_rest * cst;
}

/* ---------------------------------------------------------------------------
 * Moments, code/lb3d_analysis.F90:64-102: rho = sum(n*g), u = sum(n*g*c)
 * (u is the momentum, not the velocity).
 */
static void update_moments_site(or_state *o, double *S)
{
    for (int c = 0; c < o->p.ncomp; c++) {
        const double *n = S + OFF_N(c);
        double rho = 0.0;
        for (int s = 0; s < 19; s++) rho = rho + n[s] * (double)gdeg[s];
        S[OFF_RHO(c)] = rho;
        double u0 = 0.0, u1 = 0.0, u2 = 0.0;
        for (int s = 0; s < 19; s++) {
            double ng = n[s] * (double)gdeg[s];
            u0 = u0 + ng * (double)cx[s];
            u1 = u1 + ng * (double)cy[s];
            u2 = u2 + ng * (double)cz[s];
        }
        S[OFF_U(c)] = u0;
        S[OFF_U(c) + 1] = u1;
        S[OFF_U(c) + 2] = u2;
    }
}

/* lb3d_update_density_loop / lb3d_update_velocity_loop, code/lb3d_analysis.F90:107-153 */
static void update_moments_loop(or_state *o)
{
    #pragma omp parallel for schedule(static)
    for (int z = 0; z < o->lz; z++)
        for (int y = 0; y < o->ly; y++)
            for (int x = 0; x < o->lx; x++) update_moments_site(o, site(o, x, y, z));
}

/* ---------------------------------------------------------------------------
 * Periodic halo on one rank: whole-site copies x, then y (full x), then z
 * (full x, y); code/lb3d_mpi.F90:222-236 (the one-rank form of
 * lbe_halo_exchange, code/lb3d_bc_periodic.F90:36-56).
 */
static void copy_site(or_state *o, int xd, int yd, int zd, int xs, int ys, int zs)
{
    memcpy(site(o, xd, yd, zd), site(o, xs, ys, zs), sizeof(double) * o->sd);
}

static void halo_exchange_local(or_state *o)
{
    int nx = o->p.nx, ny = o->p.ny, nz = o->p.nz;
    for (int z = 1; z <= nz; z++)
        for (int y = 1; y <= ny; y++) {
            copy_site(o, 0, y, z, nx, y, z);
            copy_site(o, nx + 1, y, z, 1, y, z);
        }
    for (int z = 1; z <= nz; z++)
        for (int x = 0; x <= nx + 1; x++) {
            copy_site(o, x, 0, z, x, ny, z);
            copy_site(o, x, ny + 1, z, x, 1, z);
        }
    for (int y = 0; y <= ny + 1; y++)
        for (int x = 0; x <= nx + 1; x++) {
            copy_site(o, x, y, 0, x, y, nz);
            copy_site(o, x, y, nz + 1, x, y, 1);
        }
}

/* ---------------------------------------------------------------------------
 * Lees-Edwards, code/lb3d_bc_leesedwards.F90.
 * le_neighbours (:86-140), inv_fluid = 5 branch.
 */
static void le_neighbours(or_state *o)
{
    o->shear_sum = o->shear_sum + o->shear_u;
    /* lelo = mod(int(2.*shear_sum), nz): int truncates, mod takes the sign of
     * the dividend (C's % does the same) */
    long t = (long)(2.0 * o->shear_sum);
    o->lelo = (int)(t % o->p.nz);
    double lf = 2.0 * o->shear_sum;
    o->lefrac = fabs(lf - trunc(lf));
}

static inline int wrapz(int k, int n)
{
    int m = (k - 1) % n;
    if (m < 0) m += n;
    return m + 1;
}

/* Galilean correction and clamp for one LE halo site, code/lb3d_bc_leesedwards.F90:682-825
 * (bottom) and :863-1003 (top).  `S` holds the interpolated populations;
 * `base` is the site whose populations feed the clamp branch (the bottom
 * halo site itself, or N(0,y,z) for the top halo, :951-990). */
static void le_correct(or_state *o, double *S, const double *base, double two_u)
{
    int C = o->p.ncomp;
    double rN[3], p[3][3];
    for (int c = 0; c < C; c++) {
        rN[c] = 0.0;
        p[c][0] = p[c][1] = p[c][2] = 0.0;
    }
    for (int s = 0; s < 19; s++) {
        for (int c = 0; c < C; c++) {
            double n = S[OFF_N(c) + s];
            rN[c] = rN[c] + n * (double)gdeg[s];
            double ng = n * (double)gdeg[s];
            p[c][0] = p[c][0] + ng * (double)cx[s];
            p[c][1] = p[c][1] + ng * (double)cy[s];
            p[c][2] = p[c][2] + ng * (double)cz[s];
        }
    }
    double pt[3], rho_tilde = 0.0;
    for (int k = 0; k < 3; k++) pt[k] = (o->p.amass[0] * o->omega[0]) * p[0][k];
    rho_tilde = (o->p.amass[0] * rN[0]) * o->omega[0];
    for (int c = 1; c < C; c++) {
        for (int k = 0; k < 3; k++) pt[k] = pt[k] + (o->p.amass[c] * o->omega[c]) * p[c][k];
        rho_tilde = rho_tilde + (o->p.amass[c] * rN[c]) * o->omega[c];
    }
    double rho_tmp = rho_tilde > CLAMP_RHO ? rho_tilde : CLAMP_RHO;
    double ut[3] = {pt[0] / rho_tmp, pt[1] / rho_tmp, pt[2] / rho_tmp};
    double b_udu[19], b_u[19], bdu[19];
    /* u_z - shear_u - shear_u (bottom) / u_z + shear_u + shear_u (top) as two
     * separate operations, :739 and :920 */
    double uz_shift = two_u < 0 ? (ut[2] - o->shear_u) - o->shear_u
                                : (ut[2] + o->shear_u) + o->shear_u;
    or_boltz_dist(o->p.bdist, ut[0], ut[1], uz_shift, 0.0, 0.0, 0.0, b_udu);
    or_boltz_dist(o->p.bdist, ut[0], ut[1], ut[2], 0.0, 0.0, 0.0, b_u);
    for (int s = 0; s < 19; s++) bdu[s] = b_udu[s] - b_u[s];
    double df[3][19], fnew[3][19];
    int neg = 0;
    for (int c = 0; c < C; c++)
        for (int s = 0; s < 19; s++) {
            df[c][s] = rN[c] * bdu[s];
            fnew[c][s] = S[OFF_N(c) + s] + df[c][s];
            if (fnew[c][s] < 0.0) neg = 1;
        }
    if (neg) {
        int allpos = 1;
        for (int c = 0; c < C; c++)
            for (int s = 0; s < 19; s++)
                if (!(base[OFF_N(c) + s] > 0.0)) allpos = 0;
        double df_scale;
        if (allpos) {
            /* maxval(-(df)/(n)) over the concatenated (r, b, s) arrays */
            df_scale = -df[0][0] / base[OFF_N(0)];
            for (int c = 0; c < C; c++)
                for (int s = 0; s < 19; s++) {
                    double v = -df[c][s] / base[OFF_N(c) + s];
                    if (v > df_scale) df_scale = v;
                }
        } else {
            df_scale = 0.0;
        }
        for (int c = 0; c < C; c++)
            for (int s = 0; s < 19; s++)
                fnew[c][s] = fabs(df_scale) > 1.0 ? base[OFF_N(c) + s] + df[c][s] / df_scale
                                                  : base[OFF_N(c) + s];
    }
    for (int c = 0; c < C; c++)
        for (int s = 0; s < 19; s++) S[OFF_N(c) + s] = fnew[c][s];
}

/* le_halo_exchange on one rank (:147-1315).  The x swap leaves, for the
 * interior (y,z) of the halo planes (lelo >= 0, b = 0; lelo < 0, b = -1):
 *   N(0,y,z)    = (1-f) N(nx, y, z+lelo+b)   + f N(nx, y, z+lelo+1+b)
 *   N(nx+1,y,z) = f N(1, y, z-lelo-1-b)      + (1-f) N(1, y, z-lelo-b)
 * with z wrapped over 1..nz (this is the buffer algebra of :196-658 with all
 * partners equal to self), then the velocity shift of each halo site.  The y
 * and z swaps that follow are plain periodic copies of n (and da, db) over
 * the full x range, :1043-1308; the halo rows/corners of the x planes that
 * the x swap also fills are overwritten by them. */
static void le_halo_exchange(or_state *o)
{
    int nx = o->p.nx, ny = o->p.ny, nz = o->p.nz, C = o->p.ncomp;
    int lelo = o->lelo, b = lelo >= 0 ? 0 : -1;
    double f = o->lefrac;
    int nd = 19 * C + (C == 3 ? 6 : 0); /* n_r, n_b, [n_s, da, db] at offsets */
    (void)nd;
    /* bottom halo first (all y, z), then top */
    for (int y = 1; y <= ny; y++)
        for (int z = 1; z <= nz; z++) {
            double *D = site(o, 0, y, z);
            const double *A = csite(o, nx, y, wrapz(z + lelo + b, nz));
            const double *B = csite(o, nx, y, wrapz(z + lelo + 1 + b, nz));
            for (int c = 0; c < C; c++)
                for (int s = 0; s < 19; s++)
                    D[OFF_N(c) + s] = (1.0 - f) * A[OFF_N(c) + s] + f * B[OFF_N(c) + s];
            if (C == 3)
                for (int k = 0; k < 3; k++) {
                    D[OFF_DA + k] = (1.0 - f) * A[OFF_DA + k] + f * B[OFF_DA + k];
                    D[OFF_DB + k] = (1.0 - f) * A[OFF_DB + k] + f * B[OFF_DB + k];
                }
            D[off_rock(o)] = 0.0;
            le_correct(o, D, D, -1.0);
        }
    for (int y = 1; y <= ny; y++)
        for (int z = 1; z <= nz; z++) {
            double *D = site(o, nx + 1, y, z);
            const double *A = csite(o, 1, y, wrapz(z - lelo - 1 - b, nz));
            const double *B = csite(o, 1, y, wrapz(z - lelo - b, nz));
            for (int c = 0; c < C; c++)
                for (int s = 0; s < 19; s++)
                    D[OFF_N(c) + s] = f * A[OFF_N(c) + s] + (1.0 - f) * B[OFF_N(c) + s];
            if (C == 3)
                for (int k = 0; k < 3; k++) {
                    D[OFF_DA + k] = f * A[OFF_DA + k] + (1.0 - f) * B[OFF_DA + k];
                    D[OFF_DB + k] = f * A[OFF_DB + k] + (1.0 - f) * B[OFF_DB + k];
                }
            D[off_rock(o)] = 0.0;
            le_correct(o, D, csite(o, 0, y, z), +1.0);
        }
    /* y and z swaps: periodic copies over the full x (and y) range */
    for (int z = 1; z <= nz; z++)
        for (int x = 0; x <= nx + 1; x++) {
            copy_site(o, x, 0, z, x, ny, z);
            copy_site(o, x, ny + 1, z, x, 1, z);
        }
    for (int y = 0; y <= ny + 1; y++)
        for (int x = 0; x <= nx + 1; x++) {
            copy_site(o, x, y, 0, x, y, nz);
            copy_site(o, x, y, nz + 1, x, y, 1);
        }
}

/* ---------------------------------------------------------------------------
 * Boundary dispatcher, code/lb3d_bc.F90.
 */
void or_bc_init(or_state *o)
{
    /* :58-83 */
    if (o->p.lees_edwards) {
        o->shear_u = ((double)o->p.nx / ((double)o->p.nx - 1.0)) * o->shear_u;
        o->shear_tmp = o->shear_u;
        if (o->p.restore) o->shear_sum = o->shear_sum - o->shear_u;
        le_neighbours(o);
        le_halo_exchange(o);
        update_moments_loop(o);
    } else {
        halo_exchange_local(o);
    }
}

/* ---------------------------------------------------------------------------
 * Zou-He flux boundary, lbe_invade_zouandheflux (code/lb3d_invasion.F90:
 * 2143-2457), ux = uy = 0: on the post-advection populations of one component
 * at a site of the z = 1 plane (low) or the z = nz plane, the unknown
 * populations entering the lattice are set so that the z flux is uz (itype
 * < 0: uz is a mass flow, divided by the boundary density).  F = n g in the
 * reference's 1-based direction numbering.
 */
#define ZF(k) F[(k) - 1]
static void zou_he_component(double *n, double uz, int itype, int low)
{
    double F[19], rho;
    const double ux = 0.0, uy = 0.0;
    for (int s = 0; s < 19; s++) F[s] = n[s] * (double)gdeg[s];
    if (low) { /* :2179-2214 */
        if (itype < 0) {
            rho = ZF(19) + ZF(1) + ZF(2) + ZF(3) + ZF(4) + ZF(7) + ZF(11) + ZF(12) + ZF(8) +
                  2.0 * (ZF(6) + ZF(10) + ZF(14) + ZF(16) + ZF(18)) + uz;
            uz = rho > 0.0 ? uz / rho : 0.0;
        } else {
            rho = (ZF(19) + ZF(1) + ZF(2) + ZF(3) + ZF(4) + ZF(7) + ZF(11) + ZF(12) + ZF(8) +
                   2.0 * (ZF(6) + ZF(10) + ZF(14) + ZF(16) + ZF(18))) /
                  (1.0 - uz);
        }
        ZF(5) = ZF(6) + uz * rho / 3.0;
        if (itype == 0) {
            ZF(15) = ZF(18) - (ZF(3) - ZF(4)) / 4.0 + rho * uz / 6.0;
            ZF(17) = ZF(16) + (ZF(3) - ZF(4)) / 4.0 + rho * uz / 6.0;
            ZF(9) = ZF(14) - (ZF(1) - ZF(2)) / 4.0 + rho * uz / 6.0;
            ZF(13) = ZF(10) + (ZF(1) - ZF(2)) / 4.0 + rho * uz / 6.0;
        } else if (itype == 1 || itype == -1) {
            ZF(15) = ZF(18) - (ZF(7) + ZF(11) + ZF(3) - ZF(8) - ZF(12) - ZF(4)) / 2.0 + uz * rho / 6.0 +
                     uy * rho / 2.0;
            ZF(17) = ZF(16) - (ZF(8) + ZF(12) + ZF(4) - ZF(7) - ZF(11) - ZF(3)) / 2.0 + uz * rho / 6.0 -
                     uy * rho / 2.0;
            ZF(9) = ZF(14) - (ZF(7) + ZF(8) + ZF(1) - ZF(11) - ZF(12) - ZF(2)) / 2.0 + uz * rho / 6.0 +
                    ux * rho / 2.0;
            ZF(13) = ZF(10) - (ZF(11) + ZF(12) + ZF(2) - ZF(7) - ZF(8) - ZF(1)) / 2.0 + uz * rho / 6.0 -
                     ux * rho / 2.0;
        }
    } else { /* :2318-2352 */
        if (itype < 0) {
            rho = ZF(19) + ZF(1) + ZF(2) + ZF(3) + ZF(4) + ZF(7) + ZF(11) + ZF(12) + ZF(8) +
                  2.0 * (ZF(5) + ZF(9) + ZF(13) + ZF(15) + ZF(17)) - uz;
            uz = rho > 0.0 ? uz / rho : 0.0;
        } else {
            rho = (ZF(19) + ZF(1) + ZF(2) + ZF(3) + ZF(4) + ZF(7) + ZF(11) + ZF(12) + ZF(8) +
                   2.0 * (ZF(5) + ZF(9) + ZF(13) + ZF(15) + ZF(17))) /
                  (uz + 1.0);
        }
        ZF(6) = ZF(5) - uz * rho / 3.0;
        if (itype == 0) {
            ZF(18) = ZF(15) + (ZF(3) - ZF(4)) / 4.0 - rho * uz / 6.0;
            ZF(16) = ZF(17) - (ZF(3) - ZF(4)) / 4.0 - rho * uz / 6.0;
            ZF(14) = ZF(9) + (ZF(1) - ZF(2)) / 4.0 - rho * uz / 6.0;
            ZF(10) = ZF(13) - (ZF(1) - ZF(2)) / 4.0 - rho * uz / 6.0;
        } else if (itype == 1 || itype == -1) {
            ZF(18) = ZF(15) + (ZF(7) + ZF(11) + ZF(3) - ZF(8) - ZF(12) - ZF(4)) / 2.0 - uz * rho / 6.0 -
                     uy * rho / 2.0;
            ZF(16) = ZF(17) + (ZF(8) + ZF(12) + ZF(4) - ZF(7) - ZF(11) - ZF(3)) / 2.0 - uz * rho / 6.0 +
                     uy * rho / 2.0;
            ZF(14) = ZF(9) + (ZF(7) + ZF(8) + ZF(1) - ZF(11) - ZF(12) - ZF(2)) / 2.0 - uz * rho / 6.0 -
                     ux * rho / 2.0;
            ZF(10) = ZF(13) + (ZF(11) + ZF(12) + ZF(2) - ZF(7) - ZF(8) - ZF(1)) / 2.0 - uz * rho / 6.0 +
                     ux * rho / 2.0;
        }
    }
    for (int s = 0; s < 19; s++) n[s] = F[s] / (double)gdeg[s];
}
#undef ZF

void or_zou_he_site(double *n, double uz, int itype, int low) { zou_he_component(n, uz, itype, low); }

/* The boundary on both z planes of the fluid sites: surfactant first, with
 * the advected dipole of the site recomputed from the new populations and the
 * neighbours' dipole field, then red and blue.  rho and u keep the values of
 * the advection. */
static void zou_he_flux(or_state *o)
{
    const int nx = o->p.nx, ny = o->p.ny, C = o->p.ncomp, ro = off_rock(o);
    const int itype = o->p.inv_type;
    const int zs[2] = {1, o->p.nz};
    for (int b = 0; b < 2; b++) { /* nz = 1: both rules on the one plane, low first */
        const int z = zs[b], low = b == 0;
        for (int y = 1; y <= ny; y++)
            for (int x = 1; x <= nx; x++) {
                double *S = site(o, x, y, z);
                if (S[ro] != 0.0) continue;
                if (C == 3) {
                    zou_he_component(S + OFF_N(2), o->p.inv_flux[2], itype, low);
                    S[OFF_DA] = S[OFF_DA + 1] = S[OFF_DA + 2] = 0.0;
                    for (int s = 0; s < 19; s++) {
                        const double *Sb = csite(o, x - cx[s], y - cy[s], z - cz[s]);
                        for (int k = 0; k < 3; k++)
                            S[OFF_DA + k] = S[OFF_DA + k] + (S[OFF_N(2) + s] * (double)gdeg[s]) * Sb[OFF_DB + k];
                    }
                }
                zou_he_component(S + OFF_N(0), o->p.inv_flux[0], itype, low);
                zou_he_component(S + OFF_N(1), o->p.inv_flux[1], itype, low);
            }
    }
}

void or_bc_apply(or_state *o, int stage)
{
    if (stage == 6) { /* :106-122 */
        if (o->p.lees_edwards) {
            o->shear_u = o->shear_tmp * cos(o->p.shear_omega * (double)o->nt);
            le_neighbours(o);
            le_halo_exchange(o);
            update_moments_loop(o);
        } else {
            halo_exchange_local(o);
        }
    } else if (stage == 8) { /* :124-146: invasion (lbe_invade) for inv_fluid 12, then the halo */
        if (o->p.inv_fluid == 12 && !o->p.lees_edwards) zou_he_flux(o);
        if (o->p.lees_edwards) {
            le_halo_exchange(o);
            update_moments_loop(o);
        } else {
            halo_exchange_local(o);
        }
    }
}

void or_init_rock(or_state *o, const double *obs, double threshold)
{
    int nx = o->p.nx, ny = o->p.ny, nz = o->p.nz, C = o->p.ncomp;
    const int ro = off_rock(o);
    for (int z = 1; z <= nz; z++)
        for (int y = 1; y <= ny; y++)
            for (int x = 1; x <= nx; x++) {
                double *S = site(o, x, y, z);
                size_t g = (size_t)(x - 1) + (size_t)nx * ((size_t)(y - 1) + (size_t)ny * (size_t)(z - 1));
                S[ro] = obs[g] > threshold ? 1.0 : 0.0;
                if (S[ro] == 0.0) continue;
                for (int c = 0; c < C; c++)
                    for (int s = 0; s < 19; s++) S[OFF_N(c) + s] = 0.0;
                if (C == 3)
                    for (int k = 0; k < 3; k++) S[OFF_DA + k] = S[OFF_DB + k] = 0.0;
                if (o->p.rock_colour_r > 0) S[OFF_RC(0)] = o->p.rock_colour_r;
                if (o->p.rock_colour_b > 0) S[OFF_RC(1)] = o->p.rock_colour_b;
            }
}

/* The halo as the pre-advection pass builds it from the current populations,
 * without advancing the Lees-Edwards offset (le_halo_exchange with the
 * current lelo/lefrac; periodic copies otherwise). */
void or_halo_refresh(or_state *o)
{
    if (o->p.lees_edwards)
        le_halo_exchange(o);
    else
        halo_exchange_local(o);
}

/* Momentum transfer through the planes x = minx and x = maxx (global
 * 1-based), local_fluid_momentum_transfer and dump_stress,
 * code/lb3d_io_stress.F90:38-173 on one rank: vectors 9, 10 (pvecs) carry
 * +x and z momentum into plane minx from x = minx-1, their bounces 13, 14
 * leave it; mvecs 13, 14 enter plane maxx from maxx+1.  Summed over
 * components, halved; the mean shear stress is (out[1] - out[0])/(ny*nz).
 * Reads the halo as it stands (or_halo_refresh builds the pre-advection one). */
void or_stress(const or_state *o, int minx, int maxx, double out[2])
{
    static const int pvecs[2] = {8, 9}, mvecs[2] = {12, 13}; /* 0-based 9,10 and 13,14 */
    static const int bounce[19] = {1, 0, 3, 2, 5, 4, 11, 10, 13, 12, 7, 6, 9, 8, 17, 16, 15, 14, 18};
    int ny = o->p.ny, nz = o->p.nz, C = o->p.ncomp;
    double lf[2] = {0.0, 0.0};
    for (int side = 0; side < 2; side++) {
        int m = side == 0 ? minx : maxx, src = side == 0 ? m - 1 : m + 1;
        const int *vecs = side == 0 ? pvecs : mvecs;
        for (int y = 1; y <= ny; y++)
            for (int z = 1; z <= nz; z++)
                for (int t = 0; t < 2; t++) {
                    int sv = vecs[t], r = bounce[sv];
                    const double *A = csite(o, src, y, z - cz[sv]);
                    const double *B = csite(o, m, y, z);
                    for (int c = 0; c < C; c++)
                        lf[side] = lf[side] + A[OFF_N(c) + sv] * (double)cz[sv] * (double)gdeg[sv] -
                                   B[OFF_N(c) + r] * (double)cz[r] * (double)gdeg[r];
                }
    }
    out[0] = lf[0] * 0.5;
    out[1] = lf[1] * 0.5;
}

/* ---------------------------------------------------------------------------
 * Advection, code/lb3d_advection.F90:54-151.  Pull streaming of the 18 moving
 * populations from the pre-advection lattice (the rolling 3-plane buffer Nbuf
 * holds unmodified copies), the advected dipole da (ternary), then the
 * moments.  Half-way bounce-back (:85-90) is the identity without rock.
 */
void or_advection(or_state *o)
{
    int nx = o->p.nx, ny = o->p.ny, nz = o->p.nz, C = o->p.ncomp;
    const int ro = off_rock(o);
    memcpy(o->buf, o->N, sizeof(double) * (size_t)o->lx * o->ly * o->lz * o->sd);
#pragma omp parallel for schedule(static)
    for (int z = 1; z <= nz; z++)
        for (int y = 1; y <= ny; y++)
            for (int x = 1; x <= nx; x++) {
                double *S = site(o, x, y, z);
                const double *B0 = o->buf + sidx(o, x, y, z);
                if (C == 3) S[OFF_DA] = S[OFF_DA + 1] = S[OFF_DA + 2] = 0.0;
                /* half-way bounce-back, code/lb3d_bc_bounce.F90:53-90: a fluid
                 * site whose neighbour x+c_s is rock gets its own pre-advection
                 * n_s back as n_bounce(s) (ternary: plus n_s g_s db(x+c_s) in da) */
                if (B0[ro] == 0.0)
                    for (int s = 0; s < 18; s++) {
                        const double *Ba = o->buf + sidx(o, x + cx[s], y + cy[s], z + cz[s]);
                        if (Ba[ro] == 0.0) continue;
                        for (int c = 0; c < C; c++) S[OFF_N(c) + bounce_of[s]] = B0[OFF_N(c) + s];
                        if (C == 3)
                            for (int k = 0; k < 3; k++)
                                S[OFF_DA + k] = S[OFF_DA + k] + (B0[OFF_N(2) + s] * (double)gdeg[s]) * Ba[OFF_DB + k];
                    }
                /* pull from fluid neighbours into fluid sites, :118-151 */
                for (int s = 0; s < 18; s++) {
                    const double *Bs =
                        o->buf + sidx(o, x - cx[s], y - cy[s], z - cz[s]);
                    if (B0[ro] != 0.0 || Bs[ro] != 0.0) continue;
                    for (int c = 0; c < C; c++) S[OFF_N(c) + s] = Bs[OFF_N(c) + s];
                    if (C == 3) {
                        double ng = Bs[OFF_N(2) + s] * (double)gdeg[s];
                        for (int k = 0; k < 3; k++)
                            S[OFF_DA + k] = S[OFF_DA + k] + ng * Bs[OFF_DB + k];
                    }
                }
                if (C == 3) {
                    double ng = B0[OFF_N(2) + 18] * (double)gdeg[18];
                    for (int k = 0; k < 3; k++) S[OFF_DA + k] = S[OFF_DA + k] + ng * B0[OFF_DB + k];
                }
                update_moments_site(o, S);
            }
}

/* ---------------------------------------------------------------------------
 * Dipole relaxation, code/lb3d_collision.F90:296-403.
 */
static void dipole_collision(or_state *o, int x, int y, int z)
{
    const double *S0 = csite(o, x, y, z);
    double h_c[3] = {0.0, 0.0, 0.0}, h_s[3], h[3], sd_tmp[3], sd_eq[3];
    for (int k = 0; k < 3; k++) h_s[k] = S0[OFF_DA + k];
    double sN_tmp = S0[OFF_RHO(2)] > CLAMP_SURF ? S0[OFF_RHO(2)] : CLAMP_SURF;
    for (int k = 0; k < 3; k++) sd_tmp[k] = h_s[k] / sN_tmp;
    double sgn = o->p.g_bs >= 0.0 ? 1.0 : -1.0; /* sign(1.0, g_bs) */
    if (o->p.g_bs == 0.0 && signbit(o->p.g_bs)) sgn = -1.0;
    for (int s = 0; s < 18; s++) {
        const double *Sa = csite(o, x + cx[s], y + cy[s], z + cz[s]);
        double psi_r = 0.0 + Sa[OFF_RHO(0)];
        double psi_b = 0.0 + Sa[OFF_RHO(1)];
        const double *pas = Sa + OFF_DA;
        double gs = (double)gdeg[s];
        double hc_m = -((sgn * (psi_b - psi_r)) * gs);
        h_c[0] = h_c[0] + hc_m * (double)cx[s];
        h_c[1] = h_c[1] + hc_m * (double)cy[s];
        h_c[2] = h_c[2] + hc_m * (double)cz[s];
        double hs_dot = 2.0 * ((double)cx[s] * pas[0] + (double)cy[s] * pas[1] + (double)cz[s] * pas[2]);
        h_s[0] = h_s[0] + (pas[0] - (double)cx[s] * hs_dot) * gs;
        h_s[1] = h_s[1] + (pas[1] - (double)cy[s] * hs_dot) * gs;
        h_s[2] = h_s[2] + (pas[2] - (double)cz[s] * hs_dot) * gs;
    }
    for (int k = 0; k < 3; k++) h[k] = h_c[k] + h_s[k];
    double h_m = o->p.beta * sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    double third_beta = (1.0 / 3.0) * o->p.beta;
    for (int k = 0; k < 3; k++) sd_eq[k] = third_beta * h[k];
    if (h_m > 10.0e-3) {
        double exh = 1.0 / tanh(h_m);
        double sd_mag = exh - 1.0 / h_m;
        for (int k = 0; k < 3; k++) sd_eq[k] = ((o->p.beta * h[k]) / h_m) * sd_mag;
    }
    double *S = site(o, x, y, z);
    for (int k = 0; k < 3; k++) {
        sd_eq[k] = sd_eq[k] * o->p.d_0;
        S[OFF_DB + k] = o->omega_d * sd_eq[k] + (1.0 - o->omega_d) * sd_tmp[k];
    }
}

/* ---------------------------------------------------------------------------
 * Shan-Chen forces, code/lb3d_force_shanchen.F90.  psi = 1 - exp(-rho)
 * (get_psi, :334-353; psifunc has no effect without COMPAT_PSIFUN).
 */
static inline double get_psi(double rho) { return 1.0 - exp(-rho); }

/* binary, :127-210; a rock neighbour contributes through the wall couplings
 * with psi(rho + rock colour) (:167-183) */
static void sc_force_binary(const or_state *o, int x, int y, int z, double f_r[3], double f_b[3])
{
    double f_br[3] = {0, 0, 0}, f_rb[3] = {0, 0, 0}, f_wr[3] = {0, 0, 0}, f_wb[3] = {0, 0, 0};
    const int ro = off_rock(o);
    for (int s = 0; s < 18; s++) {
        const double *Sa = csite(o, x + cx[s], y + cy[s], z + cz[s]);
        const double c3[3] = {(double)cx[s], (double)cy[s], (double)cz[s]};
        if (Sa[ro] != 0.0) {
            double psi_r = get_psi(Sa[OFF_RHO(0)] + Sa[OFF_RC(0)]);
            double psi_b = get_psi(Sa[OFF_RHO(1)] + Sa[OFF_RC(1)]);
            double green_r = psi_b * (double)gdeg[s];
            double green_b = psi_r * (double)gdeg[s];
            for (int k = 0; k < 3; k++) {
                f_wb[k] = f_wb[k] + green_r * c3[k];
                f_wr[k] = f_wr[k] + green_b * c3[k];
            }
        } else {
            double psi_r = get_psi(Sa[OFF_RHO(0)]);
            double psi_b = get_psi(Sa[OFF_RHO(1)]);
            double green_r = psi_b * (double)gdeg[s];
            double green_b = psi_r * (double)gdeg[s];
            for (int k = 0; k < 3; k++) {
                f_rb[k] = f_rb[k] + green_r * c3[k];
                f_br[k] = f_br[k] + green_b * c3[k];
            }
        }
    }
    const double *S0 = csite(o, x, y, z);
    double pb = get_psi(S0[OFF_RHO(1)]), pr = get_psi(S0[OFF_RHO(0)]);
    for (int k = 0; k < 3; k++) {
        double wb = -(o->p.g_wb * f_wb[k]), wr = -(o->p.g_wr * f_wr[k]);
        f_b[k] = ((-(o->p.g_br * f_br[k])) + (wb / o->p.tau[1]) * o->p.tau_wb) * pb;
        f_r[k] = ((-(o->p.g_br * f_rb[k])) + (wr / o->p.tau[0]) * o->p.tau_wr) * pr;
    }
}

/* ternary with the dipole couplings, :215-330 */
static void sc_force_ternary(const or_state *o, int x, int y, int z, double f_r[3],
                             double f_b[3], double f_s[3])
{
    double f_br[3] = {0, 0, 0}, f_rb[3] = {0, 0, 0}, f_cs[3] = {0, 0, 0}, f_sc[3] = {0, 0, 0},
           f_ss[3] = {0, 0, 0};
    const double *S0 = csite(o, x, y, z);
    const double *dbx = S0 + OFF_DB;
    for (int s = 0; s < 18; s++) {
        const double *Sa = csite(o, x + cx[s], y + cy[s], z + cz[s]);
        double c3[3] = {(double)cx[s], (double)cy[s], (double)cz[s]};
        double gs = (double)gdeg[s];
        double psi_r = get_psi(Sa[OFF_RHO(0)]);
        double psi_b = get_psi(Sa[OFF_RHO(1)]);
        double green_r = psi_b * gs;
        double green_b = psi_r * gs;
        for (int k = 0; k < 3; k++) {
            f_rb[k] = f_rb[k] + green_r * c3[k];
            f_br[k] = f_br[k] + green_b * c3[k];
        }
        double psis = get_psi(Sa[OFF_RHO(2)]);
        double da[3];
        for (int k = 0; k < 3; k++) da[k] = Sa[OFF_DB + k] * psis;
        double d_dotc = da[0] * c3[0] + da[1] * c3[1] + da[2] * c3[2];
        for (int k = 0; k < 3; k++) f_cs[k] = f_cs[k] + (da[k] - (2.0 * d_dotc) * c3[k]) * gs;
        double di_dotc = dbx[0] * c3[0] + dbx[1] * c3[1] + dbx[2] * c3[2];
        double greenc_s = (psi_b - psi_r) * gs;
        for (int k = 0; k < 3; k++)
            f_sc[k] = f_sc[k] + greenc_s * (dbx[k] - (2.0 * di_dotc) * c3[k]);
        double di_dot_d = da[0] * dbx[0] + da[1] * dbx[1] + da[2] * dbx[2];
        double green_ss = 8.0 * gs;
        double a = di_dot_d - (2.0 * di_dotc) * d_dotc;
        for (int k = 0; k < 3; k++)
            f_ss[k] = f_ss[k] + green_ss * ((a * c3[k] + da[k] * di_dotc) + d_dotc * dbx[k]);
    }
    double pb = get_psi(S0[OFF_RHO(1)]), pr = get_psi(S0[OFF_RHO(0)]), ps = get_psi(S0[OFF_RHO(2)]);
    for (int k = 0; k < 3; k++) {
        double fbr = -(o->p.g_br * f_br[k]);
        double frb = -(o->p.g_br * f_rb[k]);
        double fsc = (2.0 * o->p.g_bs) * f_sc[k];
        double fcs = -((2.0 * o->p.g_bs) * f_cs[k]);
        double fss = -(o->p.g_ss * f_ss[k]);
        f_b[k] = (fbr + fcs) * pb;
        f_r[k] = (frb - fcs) * pr;
        f_s[k] = (fsc + fss) * ps;
    }
}

void or_sc_force(const or_state *o, int x, int y, int z, double f[3][3])
{
    if (o->p.ncomp == 3)
        sc_force_ternary(o, x, y, z, f[0], f[1], f[2]);
    else {
        sc_force_binary(o, x, y, z, f[0], f[1]);
        f[2][0] = f[2][1] = f[2][2] = 0.0;
    }
}

/* ---------------------------------------------------------------------------
 * Collision, code/lb3d_collision.F90:59-291.
 */
static int applyforce(const or_state *o, int x, int y, int z)
{
    /* lb3d_get_applyforce, :406-437, for inv_fluid not in {9, 20, 21} */
    return z >= o->p.g_accn_min[2] && z <= o->p.g_accn_max[2] && x >= o->p.g_accn_min[0] &&
           x <= o->p.g_accn_max[0] && y >= o->p.g_accn_min[1] && y <= o->p.g_accn_max[1];
}

void or_collision(or_state *o)
{
    int nx = o->p.nx, ny = o->p.ny, nz = o->p.nz, C = o->p.ncomp;
    if (C == 3) {
        /* :114-141 */
        #pragma omp parallel for schedule(static)
        for (int z = 1; z <= nz; z++)
            for (int y = 1; y <= ny; y++)
                for (int x = 1; x <= nx; x++) dipole_collision(o, x, y, z);
        halo_exchange_local(o); /* LE is rejected for ternary */
    }
    /* force loop, :151-177 */
    size_t ns = (size_t)nx * ny * nz;
    #pragma omp parallel for schedule(static)
    for (int z = 1; z <= nz; z++)
        for (int y = 1; y <= ny; y++)
            for (int x = 1; x <= nx; x++) {
                double f[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                /* only fluid sites feel the Shan-Chen force, :160-172 */
                if (csite(o, x, y, z)[off_rock(o)] == 0.0) or_sc_force(o, x, y, z, f);
                size_t i = (size_t)(x - 1) + (size_t)nx * ((size_t)(y - 1) + (size_t)ny * (size_t)(z - 1));
                for (int c = 0; c < 3; c++)
                    for (int k = 0; k < 3; k++) o->force[(c * ns + i) * 3 + k] = f[c][k];
            }
    /* collide loop, :188-289 */
    #pragma omp parallel for schedule(static)
    for (int z = 1; z <= nz; z++)
        for (int y = 1; y <= ny; y++)
            for (int x = 1; x <= nx; x++) {
                double *S = site(o, x, y, z);
                size_t i = (size_t)(x - 1) + (size_t)nx * ((size_t)(y - 1) + (size_t)ny * (size_t)(z - 1));
                double F[3][3];
                for (int c = 0; c < C; c++)
                    for (int k = 0; k < 3; k++) F[c][k] = o->force[(c * ns + i) * 3 + k];
                /* body force, lb3d_get_body_force :441-460 (z, then y, then x) */
                if (applyforce(o, x, y, z)) {
                    for (int c = 0; c < C; c++) {
                        double mass = o->p.amass[c] * S[OFF_RHO(c)];
                        F[c][2] = F[c][2] + mass * o->p.g_accn[2];
                        F[c][1] = F[c][1] + mass * o->p.g_accn[1];
                        F[c][0] = F[c][0] + mass * o->p.g_accn[0];
                    }
                }
                /* lb3d_get_sc_u_eq, :560-587 */
                double pt[3], rho_tilde, ut[3];
                for (int k = 0; k < 3; k++) pt[k] = (o->p.amass[0] * o->omega[0]) * S[OFF_U(0) + k];
                for (int c = 1; c < C; c++)
                    for (int k = 0; k < 3; k++)
                        pt[k] = pt[k] + (o->p.amass[c] * o->omega[c]) * S[OFF_U(c) + k];
                rho_tilde = (o->p.amass[0] * S[OFF_RHO(0)]) * o->omega[0];
                for (int c = 1; c < C; c++)
                    rho_tilde = rho_tilde + (o->p.amass[c] * S[OFF_RHO(c)]) * o->omega[c];
                double rt = rho_tilde > CLAMP_RHO ? rho_tilde : CLAMP_RHO;
                for (int k = 0; k < 3; k++) ut[k] = pt[k] / rt;
                for (int c = 0; c < C; c++) {
                    /* lb3d_get_sc_du, :523-556 */
                    double rho = S[OFF_RHO(c)];
                    double pon = 1.0 / (rho > CLAMP_RHO ? rho : CLAMP_RHO);
                    double du[3];
                    for (int k = 0; k < 3; k++) du[k] = ((o->p.tau[c] * pon) * F[c][k]) / o->p.amass[c];
                    /* Guo forcing term dd = (tau - 1/2) g_accn inside the body-force box
                     * (lb3d_get_guo_force, :505-519); outside it the reference leaves dd
                     * unset, here it is 0 */
                    double dd[3] = {0.0, 0.0, 0.0};
                    if (applyforce(o, x, y, z))
                        for (int k = 0; k < 3; k++) dd[k] = (o->p.tau[c] - 0.5) * o->p.g_accn[k];
                    double boltz[19];
                    or_boltz_dist_dd(o->p.bdist, ut[0], ut[1], ut[2], du[0], du[1], du[2], dd[0], dd[1], dd[2], boltz);
                    double *n = S + OFF_N(c);
                    for (int s = 0; s < 19; s++) {
                        double neq = rho * boltz[s];
                        n[s] = n[s] - o->omega[c] * (n[s] - neq);
                    }
                }
            }
}

/* ---------------------------------------------------------------------------
 * One timestep, code/lb3d_time_loop.F90:58-92 (force_apply stages are no-ops,
 * code/lb3d_force.F90:98-121).
 */
void or_time_step(or_state *o)
{
    o->nt = o->nt + 1;
    or_bc_apply(o, 6);
    or_advection(o);
    or_bc_apply(o, 8);
    or_collision(o);
}

void or_run(or_state *o, int nsteps)
{
    for (int i = 0; i < nsteps; i++) or_time_step(o);
}

/* ---------------------------------------------------------------------------
 * Sanity reductions, code/lb3d_sanity.F90:79-224 (one rank).
 */
void or_sanity(const or_state *o, double mass[3], double mom[3], int64_t counts[3])
{
    int C = o->p.ncomp;
    double nsum[3] = {0, 0, 0}, p[3] = {0, 0, 0};
    int64_t nnan = 0, nneg = 0, nfast = 0;
    for (int i = 1; i <= o->p.nx; i++)
        for (int j = 1; j <= o->p.ny; j++)
            for (int k = 1; k <= o->p.nz; k++) {
                const double *S = csite(o, i, j, k);
                if (S[off_rock(o)] != 0.0) continue; /* no_rock, :134 */
                int isnan_site = 0, isneg = 0, fast = 0;
                for (int c = 0; c < C; c++)
                    for (int s = 0; s < 19; s++) {
                        if (isnan(S[OFF_N(c) + s])) isnan_site = 1;
                        if (S[OFF_N(c) + s] < 0.0) isneg = 1;
                    }
                for (int c = 0; c < C; c++) {
                    double rho = 0.0, v[3] = {0, 0, 0};
                    for (int s = 0; s < 19; s++) rho = rho + S[OFF_N(c) + s] * (double)gdeg[s];
                    nsum[c] = nsum[c] + rho;
                    if (fabs(rho) > 1.0e-12) {
                        double m[3] = {0, 0, 0};
                        for (int s = 0; s < 19; s++) {
                            double ng = S[OFF_N(c) + s] * (double)gdeg[s];
                            m[0] = m[0] + ng * (double)cx[s];
                            m[1] = m[1] + ng * (double)cy[s];
                            m[2] = m[2] + ng * (double)cz[s];
                        }
                        for (int a = 0; a < 3; a++) v[a] = m[a] * o->p.amass[c] / rho;
                        if (sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) > 0.1) fast = 1;
                    }
                }
                /* massflow: sum over components of amass * sum(n*g*c) */
                for (int c = 0; c < C; c++) {
                    double m[3] = {0, 0, 0};
                    for (int s = 0; s < 19; s++) {
                        double ng = S[OFF_N(c) + s] * (double)gdeg[s];
                        m[0] = m[0] + ng * (double)cx[s];
                        m[1] = m[1] + ng * (double)cy[s];
                        m[2] = m[2] + ng * (double)cz[s];
                    }
                    for (int a = 0; a < 3; a++) p[a] = p[a] + o->p.amass[c] * m[a];
                }
                nnan += isnan_site;
                nneg += isneg;
                nfast += fast;
            }
    for (int c = 0; c < 3; c++) mass[c] = nsum[c];
    for (int a = 0; a < 3; a++) mom[a] = p[a];
    counts[0] = nnan;
    counts[1] = nneg;
    counts[2] = nfast;
}

/* The full sanity check, code/lb3d_sanity.F90:79-384 on one rank, in its loop
 * order (x outermost, z innermost): per species the velocity of each site
 * with |rho| > 1e-12, its running sum, and the largest |component| with the
 * site of the last `>=` update; totals and site counts as or_sanity. */
void or_sanity_full(const or_state *o, or_sanity_report *r)
{
    int C = o->p.ncomp;
    memset(r, 0, sizeof *r);
    double maxvel[3][3] = {{0}};
    for (int i = 1; i <= o->p.nx; i++)
        for (int j = 1; j <= o->p.ny; j++)
            for (int k = 1; k <= o->p.nz; k++) {
                const double *S = csite(o, i, j, k);
                if (S[off_rock(o)] != 0.0) continue; /* no_rock, :134 */
                r->fluid_sites += 1;
                int isnan_site = 0, isneg = 0, fast = 0;
                for (int c = 0; c < C; c++)
                    for (int s = 0; s < 19; s++) {
                        if (isnan(S[OFF_N(c) + s])) isnan_site = 1;
                        if (S[OFF_N(c) + s] < 0.0) isneg = 1;
                    }
                for (int c = 0; c < C; c++) {
                    double rho = 0.0;
                    for (int s = 0; s < 19; s++) rho = rho + S[OFF_N(c) + s] * (double)gdeg[s];
                    r->mass[c] = r->mass[c] + rho;
                    if (fabs(rho) > 1.0e-12) {
                        double m[3] = {0, 0, 0}, v[3];
                        for (int s = 0; s < 19; s++) {
                            double ng = S[OFF_N(c) + s] * (double)gdeg[s];
                            m[0] = m[0] + ng * (double)cx[s];
                            m[1] = m[1] + ng * (double)cy[s];
                            m[2] = m[2] + ng * (double)cz[s];
                        }
                        for (int a = 0; a < 3; a++) {
                            v[a] = m[a] * o->p.amass[c] / rho;
                            r->vel_sum[c][a] = r->vel_sum[c][a] + v[a];
                            if (fabs(v[a]) >= maxvel[c][a]) {
                                maxvel[c][a] = fabs(v[a]);
                                r->max_pos[c][a][0] = i;
                                r->max_pos[c][a][1] = j;
                                r->max_pos[c][a][2] = k;
                            }
                        }
                        if (sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) > 0.1) fast = 1;
                    }
                }
                for (int c = 0; c < C; c++) {
                    double m[3] = {0, 0, 0};
                    for (int s = 0; s < 19; s++) {
                        double ng = S[OFF_N(c) + s] * (double)gdeg[s];
                        m[0] = m[0] + ng * (double)cx[s];
                        m[1] = m[1] + ng * (double)cy[s];
                        m[2] = m[2] + ng * (double)cz[s];
                    }
                    for (int a = 0; a < 3; a++) r->momentum[a] = r->momentum[a] + o->p.amass[c] * m[a];
                }
                r->nan_sites += isnan_site;
                r->negative_sites += isneg;
                r->fast_sites += fast;
            }
    for (int c = 0; c < 3; c++)
        for (int a = 0; a < 3; a++) r->max_vel[c][a] = maxvel[c][a];
}

/* ---------------------------------------------------------------------------
 * Dump observables, code/lb3d_io_dump_data.F90, interior, x fastest.
 */
static double rho_of(const double *S, int c)
{
    double r = 0.0;
    for (int s = 0; s < 19; s++) r = r + S[OFF_N(c) + s] * (double)gdeg[s];
    return r;
}
static double mom_of(const double *S, int c, const int *cv)
{
    double r = 0.0;
    for (int s = 0; s < 19; s++) r = r + (S[OFF_N(c) + s] * (double)gdeg[s]) * (double)cv[s];
    return r;
}

#define FOR_INTERIOR(o)                          \
    size_t _i = 0;                               \
    for (int z = 1; z <= (o)->p.nz; z++)         \
        for (int y = 1; y <= (o)->p.ny; y++)     \
            for (int x = 1; x <= (o)->p.nx; x++, _i++)

void or_dump_od(const or_state *o, double *out)
{ /* dump_od, :1100-1130 */
    FOR_INTERIOR(o) out[_i] = csite(o, x, y, z)[off_rock(o)] == 0.0 ? o->p.amass[0] * rho_of(csite(o, x, y, z), 0) : 0.0;
}
void or_dump_wd(const or_state *o, double *out)
{ /* dump_wd, :1136-1166 */
    FOR_INTERIOR(o) out[_i] = csite(o, x, y, z)[off_rock(o)] == 0.0 ? o->p.amass[1] * rho_of(csite(o, x, y, z), 1) : 0.0;
}
void or_dump_sur(const or_state *o, double *out)
{ /* dump_sur, :922-948 */
    FOR_INTERIOR(o) out[_i] = o->p.ncomp == 3 ? o->p.amass[2] * rho_of(csite(o, x, y, z), 2) : 0.0;
}
void or_dump_dir(const or_state *o, double *out)
{ /* dump_dir, :955-980 */
    FOR_INTERIOR(o)
    {
        const double *S = csite(o, x, y, z);
        for (int k = 0; k < 3; k++) out[3 * _i + k] = o->p.ncomp == 3 ? S[OFF_DB + k] : 0.0;
    }
}
void or_dump_vel(const or_state *o, double *out)
{ /* dump_vel, :986-1096 (SCMP false) */
    int C = o->p.ncomp;
    FOR_INTERIOR(o)
    {
        const double *S = csite(o, x, y, z);
        if (S[off_rock(o)] != 0.0) { /* rock sites dump 0, :1031 */
            out[_i] = 0.0;
            continue;
        }
        double v = (mom_of(S, 0, cz) * o->p.amass[0]) * o->omega[0];
        double den = (rho_of(S, 0) * o->p.amass[0]) * o->omega[0];
        for (int c = 1; c < C; c++) {
            v = v + (mom_of(S, c, cz) * o->p.amass[c]) * o->omega[c];
            den = den + (rho_of(S, c) * o->p.amass[c]) * o->omega[c];
        }
        v = v / (den > 10.0e-9 ? den : 10.0e-9);
        if (applyforce(o, x, y, z)) v = v - o->p.g_accn[2] / 2.0;
        out[_i] = v;
    }
}
void or_dump_arr(const or_state *o0, double *out)
{ /* dump_arrows, :1308-1460: (sum_c amass_c (P_c - F_c/2 rho_c)) / sum_c amass_c rho_c.
   * The XDR dumps run on the gathered lattice Nm, whose halo is refilled
   * periodically (postprocess, code/lb3d_io.F90:387-442, recv_final_lattice,
   * code/lb3d_mpi.F90:788-858), so the forces here see periodic halos also
   * under Lees-Edwards. */
    int C = o0->p.ncomp;
    const int *cv[3] = {cx, cy, cz};
    or_state tmp = *o0;
    size_t bytes = sizeof(double) * (size_t)o0->lx * o0->ly * o0->lz * o0->sd;
    tmp.N = (double *)malloc(bytes);
    if (!tmp.N) return;
    memcpy(tmp.N, o0->N, bytes);
    halo_exchange_local(&tmp);
    const or_state *o = &tmp;
    FOR_INTERIOR(o)
    {
        const double *S = csite(o, x, y, z);
        const int fluid = S[off_rock(o)] == 0.0;
        double f[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        if (fluid) or_sc_force(o, x, y, z, f);
        double den = rho_of(S, 0) * o->p.amass[0];
        for (int c = 1; c < C; c++) den = den + rho_of(S, c) * o->p.amass[c];
        /* max(10.e-9, real(...)): single-precision literal and conversion */
        float denf = (float)den;
        double dd = (double)denf > (double)10.e-9f ? (double)denf : (double)10.e-9f;
        for (int k = 0; k < 3; k++) {
            /* (((A_r - B_r) + A_b) - B_b) + ... left to right */
            double v = 0.0;
            for (int c = 0; c < C; c++) {
                double a = mom_of(S, c, cv[k]) * o->p.amass[c];
                double b = ((f[c][k] / 2.0) * rho_of(S, c)) * o->p.amass[c];
                v = c == 0 ? a - b : (v + a) - b;
            }
            v = v / dd;
            /* external force term inside the body-force box, :1428-1442 */
            if (applyforce(o, x, y, z) && fluid) v = v - o->p.g_accn[k] / 2.0;
            out[3 * _i + k] = v;
        }
    }
    free(tmp.N);
}

/* ---------------------------------------------------------------------------
 * Seeded synthetic initial state (not a reference routine: the configs'
 * inputs are generated identically for the oracle and the GPU library).
 */
static uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double or_hash_uniform(uint64_t seed, uint32_t stream, uint64_t s)
{
    uint64_t h = mix64(mix64(seed ^ ((uint64_t)stream << 48)) ^ s);
    return (double)(h >> 11) * 0x1.0p-53;
}

void or_init_noise(or_state *o, uint64_t seed, const double rho_mean[3], const double amp[3],
                   int dipoles)
{
    int nx = o->p.nx, ny = o->p.ny, nz = o->p.nz, C = o->p.ncomp;
    const double w_move = 1.0 / 36.0, w_rest = 1.0 / 3.0;
    #pragma omp parallel for schedule(static)
    for (int z = 1; z <= nz; z++)
        for (int y = 1; y <= ny; y++)
            for (int x = 1; x <= nx; x++) {
                uint64_t g = (uint64_t)(x - 1) + (uint64_t)nx * ((uint64_t)(y - 1) + (uint64_t)ny * (uint64_t)(z - 1));
                double *S = site(o, x, y, z);
                for (int c = 0; c < C; c++) {
                    double u = or_hash_uniform(seed, (uint32_t)c, g);
                    double rho = rho_mean[c] + amp[c] * (2.0 * u - 1.0);
                    for (int s = 0; s < 18; s++) S[OFF_N(c) + s] = w_move * rho;
                    S[OFF_N(c) + 18] = w_rest * rho;
                }
                if (C == 3) {
                    for (int k = 0; k < 3; k++) S[OFF_DA + k] = 0.0;
                    if (dipoles) {
                        double zc = 2.0 * or_hash_uniform(seed, 3u, g) - 1.0;
                        double ph = 6.283185307179586 * or_hash_uniform(seed, 4u, g);
                        double r = sqrt(1.0 - zc * zc);
                        S[OFF_DB] = r * cos(ph);
                        S[OFF_DB + 1] = r * sin(ph);
                        S[OFF_DB + 2] = zc;
                    } else {
                        S[OFF_DB] = S[OFF_DB + 1] = S[OFF_DB + 2] = 0.0;
                    }
                }
                S[off_rock(o)] = 0.0;
                update_moments_site(o, S);
            }
}
