// lb3d_lattice.h -- D3Q19 constants, the equilibrium distribution and the
// kernel-side parameter block shared by all LB3D kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lb3d {

// D3Q19 vectors (index 18 = rest) and the D3Q25 projection degeneracies g,
// code/lb3d_lattice.F90:137-164.  g also serves as the Shan-Chen stencil
// weight g_sc (code/lb3d_force_shanchen.F90:145-149).
__device__ constexpr int CX[19] = {1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
__device__ constexpr int CY[19] = {0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0};
__device__ constexpr int CZ[19] = {0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, 0};
__device__ constexpr int GD[19] = {2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
// opposite vector (code/lb3d_lattice.F90:173-177, 0-based)
__device__ constexpr int BOUNCE[19] = {1, 0, 3, 2, 5, 4, 11, 10, 13, 12, 7, 6, 9, 8, 17, 16, 15, 14, 18};

// Host copies of the same tables.
namespace host {
constexpr int CX[19] = {1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0, 0};
constexpr int CY[19] = {0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0};
constexpr int CZ[19] = {0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, 0};
constexpr int GD[19] = {2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
}  // namespace host

// Float literals of the reference widened to double: dble(10.e-9) and
// dble(10.e-5) (code/lb3d_collision.F90:345, 546, 586).
constexpr double CLAMP_RHO = 9.99999993922529e-09;
constexpr double CLAMP_SURF = 9.999999747378752e-05;

// Kernel-side parameters, converted once on the host into the working type R.
template <typename R>
struct Phys {
    int nx, ny, nz;
    uint32_t ns;          // sites nx*ny*nz
    int bdist;
    int body;             // body force box active
    int bmin[3], bmax[3]; // 1-based inclusive box (x, y, z)
    R amass[3], tau[3], omega[3], am_om[3];  // am_om = amass*omega
    R tau_am[3];                             // tau/amass (the FP32 form's force-to-velocity factor)
    R omega_d, one_m_omega_d, beta, third_beta, d_0;
    R g_br, g_bs, g_ss, two_g_bs, sign_g_bs;
    R g_accn[3];
    // deviation storage: w_i*rho0 per component and psi(rho0), exp(-rho0)
    R rho0[3], wrho0_move[3], wrho0_rest[3], psi0[3], emr0[3];
    // walls: g_wr, g_wb, tau_wr/tau_r, tau_wb/tau_b as the force uses them, rock colours
    R g_wr, g_wb, tau_wr, tau_wb, rock_col[2];
    // Zou-He flux boundary on the z planes: on, inv_type, z flux per component
    int zh, zh_type;
    R zh_flux[3];
};

// Lees-Edwards scalars, updated on the device once per step (FP64).
struct LEState {
    double shear_u, shear_tmp, shear_sum, lefrac, shear_omega;
    int lelo;
    int nt;
};
// adv argument of the Lees-Edwards halo-plane kernels and of the collide's
// state update: the stored state is the step's own (k_le_neighbours has
// advanced it); any other value is the nt_set of le_advance
constexpr int LE_STATE_AS_IS = -2;

// The other equilibrium forms of code/lb3d_bdist.F90 (COMPAT_BDIST): 3 second
// order in (u, du) (:372-391), 4 and 5 Guo's second and fourth order with the
// forcing term dd (:392-455), 6 the fourth-order Luo form (:464-482), 1 the
// iterated exponential (Hudong) form (:237-300).  Integer powers are products,
// x**4 as (x*x)*(x*x); divisions by T are divisions, as written.
template <typename R>
__device__ __forceinline__ void boltz_general(int bdist, R ux, R uy, R uz, R dux, R duy, R duz, R ddx, R ddy,
                                              R ddz, R b[19])
{
    const R T = R(1.0 / 3.0), T2 = T * T, T3 = T2 * T;
    const R w = R(1.0 / 36.0), w0 = R(1.0 / 3.0);
    if (bdist == 3) {
        R uke = R(0.5) * (ux * ux + uy * uy + uz * uz);
        R udotdu = ux * dux + uy * duy + uz * duz;
        for (int s = 0; s < 18; s++) {
            R cu = R(CX[s]) * ux + R(CY[s]) * uy + R(CZ[s]) * uz;
            R cdu = R(CX[s]) * dux + R(CY[s]) * duy + R(CZ[s]) * duz;
            R q = cu / T;
            b[s] = w * (R(1.0) + q + R(0.5) * (q * q) - uke / T + cdu / T + (cu * cdu) / T2 - udotdu / T);
        }
        b[18] = w0 * (R(1.0) - uke / T - udotdu / T);
        return;
    }
    if (bdist == 4 || bdist == 5 || bdist == 6) {
        R tux = ux + dux, tuy = uy + duy, tuz = uz + duz;
        R uke = R(0.5) * (tux * tux + tuy * tuy + tuz * tuz);
        R dke = bdist == 6 ? R(0.0) : R(0.5) * (ddx * ddx + ddy * ddy + ddz * ddz);
        R t1x = bdist == 5 ? tux - ddx : tux, t1y = bdist == 5 ? tuy - ddy : tuy, t1z = bdist == 5 ? tuz - ddz : tuz;
        R u1ke = R(0.5) * (t1x * t1x + t1y * t1y + t1z * t1z);
        for (int s = 0; s < 18; s++) {
            R cu = R(CX[s]) * tux + R(CY[s]) * tuy + R(CZ[s]) * tuz;
            R q = cu / T;
            if (bdist == 6) {
                R q2 = q * q, h = uke / T;
                b[s] = w * (R(1.0) + q + R(0.5) * q2 - uke / T - (uke * cu) / T2 + (q2 * q) / R(6.0) +
                            R(0.5) * (h * h) - (R(0.5) * ((cu * cu) * uke)) / T3 + (q2 * q2) / R(24.0));
                continue;
            }
            R cd = R(CX[s]) * ddx + R(CY[s]) * ddy + R(CZ[s]) * ddz;
            R r = cd / T;
            R v = R(1.0) + q + R(0.5) * (q * q - r * r) - (uke - dke) / T;
            if (bdist == 5) {
                R c1 = R(CX[s]) * t1x + R(CY[s]) * t1y + R(CZ[s]) * t1z;
                R q1 = c1 / T, q12 = q1 * q1, h1 = u1ke / T;
                v = v - (uke * c1) / T2 + (q12 * q1) / R(6.0) + R(0.5) * (h1 * h1) -
                    (R(0.5) * ((c1 * c1) * u1ke)) / T3 + (q12 * q12) / R(24.0);
            }
            b[s] = w * v;
        }
        R h = uke / T;
        if (bdist == 4)
            b[18] = w0 * (R(1.0) - (uke - dke) / T);
        else if (bdist == 5)
            b[18] = w0 * (R(1.0) - (uke - dke) / T + R(0.5) * (h * h));
        else
            b[18] = w0 * (R(1.0) - uke / T + R(0.5) * (h * h));
        return;
    }
    // bdist == 1
    R uxf = ux + dux, uyf = uy + duy, uzf = uz + duz;
    R uke_eq = R(0.5) * (uxf * uxf + uyf * uyf + uzf * uzf);
    R u_mag = sqrt(R(2.0) * uke_eq);
    int ifactor = (int)(R(200.0) * u_mag);
    R n_iter = R(ifactor > 2 ? ifactor : 2);
    R delta = sqrt(R(1.0) / n_iter);
    R uxe = uxf, uye = uyf, uze = uzf, rho_eq = R(1.0);
    for (int it = 1; it <= (int)n_iter; it++) {
        R ex[3] = {exp(uxe / T), exp(uye / T), exp(uze / T)};
        R expek = R(1.0) / exp(uke_eq / T);
        for (int s = 0; s < 18; s++) {
            const int cv[3] = {CX[s], CY[s], CZ[s]};
            R a = R(1.0);
            for (int k = 0; k < 3; k++) {
                R pk = cv[k] == 1 ? ex[k] : (cv[k] == -1 ? R(1.0) / ex[k] : R(1.0));
                a = k == 0 ? pk : a * pk;
            }
            b[s] = ((rho_eq * w) * a) * expek;
        }
        b[18] = (rho_eq * w0) * expek;
        R r1 = R(0.0), px = R(0.0), py = R(0.0), pz = R(0.0);
        for (int s = 0; s < 18; s++) {
            r1 = r1 + b[s];
            px = px + b[s] * R(CX[s]);
            py = py + b[s] * R(CY[s]);
            pz = pz + b[s] * R(CZ[s]);
        }
        R rho_tt = r1 + b[18];
        rho_eq = rho_eq / (R(1.0) + (delta * (rho_tt - R(1.0))) / rho_eq);
        uxe = uxe + delta * (uxf - px);
        uye = uye + delta * (uyf - py);
        uze = uze + delta * (uzf - pz);
        uke_eq = R(0.5) * (uxe * uxe + uye * uye + uze * uze);
    }
}

// Equilibrium distribution, code/lb3d_bdist.F90:175-370 (COMPAT_BDIST):
// bdist 0 = Luo (:214-234), 2 = third-order Taylor (:301-370), in the
// reference's operation order.  b[i] are the weights w_i * (...), so that
// the equilibrium population is rho*b[i].
template <typename R>
__device__ __forceinline__ void boltz_dist(int bdist, R ux, R uy, R uz, R dux, R duy, R duz,
                                           R b[19])
{
    const R holdit = R(3.0);  // 1/T with T = 1/3
    const R w = R(1.0 / 36.0);
    if (bdist != 0 && bdist != 2) {
        boltz_general<R>(bdist, ux, uy, uz, dux, duy, duz, R(0), R(0), R(0), b);
        return;
    }
    if (bdist == 0) {
        R uke = R(0.5) * (ux * ux + uy * uy + uz * uz);
        R udotdu = ux * dux + uy * duy + uz * duz;
        R uholdit = uke * holdit;
#pragma unroll
        for (int s = 0; s < 18; s++) {
            // c.u over the nonzero components of c (a zero term only changes the
            // sign of an exact zero)
            R cdotu = R(0), cdotdu = R(0);
            if (CX[s] != 0) { cdotu = cdotu + R(CX[s]) * ux; cdotdu = cdotdu + R(CX[s]) * dux; }
            if (CY[s] != 0) { cdotu = cdotu + R(CY[s]) * uy; cdotdu = cdotdu + R(CY[s]) * duy; }
            if (CZ[s] != 0) { cdotu = cdotu + R(CZ[s]) * uz; cdotdu = cdotdu + R(CZ[s]) * duz; }
            R ch = cdotu * holdit;
            R cdh = cdotdu * holdit;
            R ch2 = ch * ch;
            b[s] = w * (R(1.0) + ch + R(0.5) * ch2 - uholdit + (ch * ch2) / R(6.0) - uholdit * ch +
                        cdh + ch * cdh - udotdu * holdit + R(0.5) * ch2 * cdh -
                        ch * udotdu * holdit - uholdit * cdh);
        }
        const R T = R(1.0 / 3.0);
        b[18] = R(1.0 / 3.0) * (R(1.0) - uke / T - udotdu / T);
        return;
    }
    R tux = ux + dux, tuy = uy + duy, tuz = uz + duz;
    R tx = tux * holdit, ty = tuy * holdit, tz = tuz * holdit;
    R uke = R(0.5) * (tux * tx + tuy * ty + tuz * tz);
    R cst = R(1.0) - uke;
    R c1 = w * cst;
    const R wby2 = w / R(2.0), wby6 = w / R(6.0);
    auto pair = [&](R cu, int ip, int im) {
        R even = c1 + wby2 * (cu * cu);
        R odd = c1 * cu + wby6 * (cu * cu * cu);
        b[ip] = even + odd;
        b[im] = even - odd;
    };
    pair(tx, 0, 1);
    pair(ty, 2, 3);
    pair(tz, 4, 5);
    pair(tx + ty, 6, 11);
    pair(tx - ty, 7, 10);
    pair(tx + tz, 8, 13);
    pair(tx - tz, 9, 12);
    pair(ty + tz, 14, 17);
    pair(ty - tz, 15, 16);
    b[18] = R(1.0 / 3.0) * cst;
}

// Deviation form of the equilibrium for FP32 storage: bw[i] = b[i] - w_i,
// computed without forming b (only small terms), bdist 2 and 0.
__device__ __forceinline__ void boltz_dev(int bdist, float ux, float uy, float uz, float dux,
                                          float duy, float duz, float bw[19])
{
    const float w = 1.0f / 36.0f;
    if (bdist == 0) {
        float uke = 0.5f * (ux * ux + uy * uy + uz * uz);
        float udotdu = ux * dux + uy * duy + uz * duz;
        float uh = 3.0f * uke;
#pragma unroll
        for (int s = 0; s < 18; s++) {
            float cu = 0.0f, cd = 0.0f;
            if (CX[s] != 0) { cu = cu + float(CX[s]) * ux; cd = cd + float(CX[s]) * dux; }
            if (CY[s] != 0) { cu = cu + float(CY[s]) * uy; cd = cd + float(CY[s]) * duy; }
            if (CZ[s] != 0) { cu = cu + float(CZ[s]) * uz; cd = cd + float(CZ[s]) * duz; }
            float ch = 3.0f * cu;
            float cdh = 3.0f * cd;
            float ch2 = ch * ch;
            bw[s] = w * (ch + 0.5f * ch2 - uh + (ch * ch2) * (1.0f / 6.0f) - uh * ch + cdh +
                         ch * cdh - 3.0f * udotdu + 0.5f * ch2 * cdh - 3.0f * ch * udotdu - uh * cdh);
        }
        bw[18] = (1.0f / 3.0f) * (-3.0f * uke - 3.0f * udotdu);
        return;
    }
    float tx = 3.0f * (ux + dux), ty = 3.0f * (uy + duy), tz = 3.0f * (uz + duz);
    float uke = (1.0f / 6.0f) * (tx * tx + ty * ty + tz * tz);
    float wc = w * (1.0f - uke);
    float wuke = -w * uke;
    auto pair = [&](float cu, int ip, int im) {
        float even = wuke + (0.5f * w) * (cu * cu);
        float odd = wc * cu + (w / 6.0f) * (cu * cu * cu);
        bw[ip] = even + odd;
        bw[im] = even - odd;
    };
    pair(tx, 0, 1);
    pair(ty, 2, 3);
    pair(tz, 4, 5);
    pair(tx + ty, 6, 11);
    pair(tx - ty, 7, 10);
    pair(tx + tz, 8, 13);
    pair(tx - tz, 9, 12);
    pair(ty + tz, 14, 17);
    pair(ty - tz, 15, 16);
    bw[18] = -(1.0f / 3.0f) * uke;
}

__device__ __forceinline__ int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

}  // namespace lb3d
