// lb3d_launch.h -- the Launcher implementation, included by exactly one
// translation unit per storage form (lb3d_fp64.hip, lb3d_fp32.hip).
#pragma once

#include <type_traits>

#include <cmath>
#include <cstring>

#include "lb3d_kernels.h"
#include "lb3d_solver.h"

namespace lb3d {

template <typename R>
Phys<R> make_phys(const lb3d_params &p)
{
    Phys<R> P{};
    P.nx = p.nx;
    P.ny = p.ny;
    P.nz = p.nz;
    P.ns = (uint32_t)p.nx * (uint32_t)p.ny * (uint32_t)p.nz;
    P.bdist = p.bdist;
    P.body = (p.g_accn[0] != 0.0 || p.g_accn[1] != 0.0 || p.g_accn[2] != 0.0) ? 1 : 0;
    const int ext[3] = {p.nx, p.ny, p.nz};
    for (int k = 0; k < 3; k++) {
        P.bmin[k] = p.g_accn_min[k];
        P.bmax[k] = p.g_accn_max[k] == 0 ? ext[k] : p.g_accn_max[k];
    }
    for (int c = 0; c < 3; c++) {
        double om = 1.0 / p.tau[c];
        P.amass[c] = R(p.amass[c]);
        P.tau[c] = R(p.tau[c]);
        P.omega[c] = R(om);
        P.am_om[c] = R(p.amass[c] * om);
        P.tau_am[c] = R(p.tau[c] / p.amass[c]);
        P.g_accn[c] = R(p.g_accn[c]);
        double r0 = p.rho0[c] > 0.0 ? p.rho0[c] : 0.7;
        P.rho0[c] = R(r0);
        P.wrho0_move[c] = R(r0 / 36.0);
        P.wrho0_rest[c] = R(r0 / 3.0);
        P.psi0[c] = R(1.0 - std::exp(-r0));
        P.emr0[c] = R(std::exp(-r0));
    }
    double omd = 1.0 / p.tau_d;
    P.omega_d = R(omd);
    P.one_m_omega_d = R(1.0 - omd);
    P.beta = R(p.beta);
    P.third_beta = R((1.0 / 3.0) * p.beta);
    P.d_0 = R(p.d_0);
    P.g_br = R(p.g_br);
    P.g_bs = R(p.g_bs);
    P.g_ss = R(p.g_ss);
    P.two_g_bs = R(2.0 * p.g_bs);
    P.sign_g_bs = R(std::signbit(p.g_bs) ? -1.0 : 1.0);
    P.g_wr = R(p.g_wr);
    P.g_wb = R(p.g_wb);
    P.tau_wr = R(p.tau_wr);
    P.tau_wb = R(p.tau_wb);
    P.rock_col[0] = R(p.rock_colour_r);
    P.rock_col[1] = R(p.rock_colour_b);
    P.zh = p.inv_fluid == 12 ? 1 : 0;
    P.zh_type = p.inv_type;
    P.zh_flux[0] = R(p.inv_flux[0]);
    P.zh_flux[1] = R(p.inv_flux[1]);
    P.zh_flux[2] = R(p.inv_flux[2]);
    return P;
}

// Seeded synthetic state (shares the hash of the oracle's or_init_noise).
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double hash_uniform(uint64_t seed, uint32_t stream, uint64_t s)
{
    uint64_t h = mix64(mix64(seed ^ ((uint64_t)stream << 48)) ^ s);
    return (double)(h >> 11) * 0x1.0p-53;
}

// Wavefront reductions for the diagnostics: a butterfly over the 64 lanes
// (cross-lane permutes, no LDS), after which every lane holds the result;
// the block then combines its waves' results in wave order.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
    return v;
}
// the largest m over the wave; on equal m the larger key (the site that
// comes later in the reference's loop), the same pair on every lane
__device__ __forceinline__ void wave_argmax(double &m, double &key)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double m1 = __shfl_xor(m, o, 64), k1 = __shfl_xor(key, o, 64);
        if (m1 > m || (m1 == m && k1 > key)) {
            m = m1;
            key = k1;
        }
    }
}
constexpr int RED_WAVES = 256 / 64;  // waves of a 256-thread diagnostics block

template <typename R, bool Dev, int C>
__global__ __launch_bounds__(256) void k_init_noise(Phys<R> P, uint64_t seed, double m0, double m1,
                                                    double m2, double a0, double a1, double a2,
                                                    double r00, double r01, double r02, int dipoles,
                                                    R *__restrict__ f, R *__restrict__ db)
{
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.ns) return;
    const double mean[3] = {m0, m1, m2}, amp[3] = {a0, a1, a2}, r0[3] = {r00, r01, r02};
#pragma unroll
    for (int c = 0; c < C; c++) {
        double u = hash_uniform(seed, (uint32_t)c, (uint64_t)s);
        double rho = mean[c] + amp[c] * (2.0 * u - 1.0);
        R mv = Dev ? R((1.0 / 36.0) * (rho - r0[c])) : R((1.0 / 36.0) * rho);
        R rv = Dev ? R((1.0 / 3.0) * (rho - r0[c])) : R((1.0 / 3.0) * rho);
        for (int i = 0; i < 18; i++) f[((uint32_t)c * 19u + (uint32_t)i) * P.ns + s] = mv;
        f[((uint32_t)c * 19u + 18u) * P.ns + s] = rv;
    }
    if (C == 3) {
        double d0 = 0.0, d1 = 0.0, d2 = 0.0;
        if (dipoles) {
            double zc = 2.0 * hash_uniform(seed, 3u, (uint64_t)s) - 1.0;
            double ph = 6.283185307179586 * hash_uniform(seed, 4u, (uint64_t)s);
            double r = sqrt(1.0 - zc * zc);
            d0 = r * cos(ph);
            d1 = r * sin(ph);
            d2 = zc;
        }
        db[s] = R(d0);
        db[P.ns + s] = R(d1);
        db[2u * P.ns + s] = R(d2);
    }
}

// Sanity reductions (code/lb3d_sanity.F90:79-224) on the post-collision
// populations: per-block FP64 partial sums, reduced on the host.
template <typename R, bool Dev, int C>
__global__ __launch_bounds__(256) void k_diagnostics(Phys<R> P, const R *__restrict__ f,
                                                     const uint8_t *__restrict__ rock,
                                                     double *__restrict__ partials)
{
    __shared__ double red[NDIAG][RED_WAVES];
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    double v[NDIAG];
    for (int k = 0; k < NDIAG; k++) v[k] = 0.0;
    if (s < P.ns && !(rock && rock[s])) {  // fluid sites only
        bool isnan_ = false, isneg = false, fast = false;
        for (int c = 0; c < C; c++) {
            double rho = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0;
            for (int i = 0; i < 19; i++) {
                double n = (double)f[((uint32_t)c * 19u + (uint32_t)i) * P.ns + s];
                if (Dev) n = (i < 18 ? (double)P.rho0[c] / 36.0 : (double)P.rho0[c] / 3.0) + n;
                if (n != n) isnan_ = true;
                if (n < 0.0) isneg = true;
                double ng = n * (double)GD[i];
                rho = rho + ng;
                m0 = m0 + ng * (double)CX[i];
                m1 = m1 + ng * (double)CY[i];
                m2 = m2 + ng * (double)CZ[i];
            }
            v[c] = rho;
            double am = (double)P.amass[c];
            v[3] = v[3] + am * m0;
            v[4] = v[4] + am * m1;
            v[5] = v[5] + am * m2;
            if (fabs(rho) > 1.0e-12) {
                double a0 = m0 * am / rho, a1 = m1 * am / rho, a2 = m2 * am / rho;
                if (sqrt(a0 * a0 + a1 * a1 + a2 * a2) > 0.1) fast = true;
            }
        }
        v[6] = isnan_ ? 1.0 : 0.0;
        v[7] = isneg ? 1.0 : 0.0;
        v[8] = fast ? 1.0 : 0.0;
    }
    const unsigned wv = threadIdx.x / 64u, ln = threadIdx.x % 64u;
    for (int k = 0; k < NDIAG; k++) {
        const double t = wave_sum(v[k]);
        if (ln == 0) red[k][wv] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = 0; k < NDIAG; k++) {
            double t = red[k][0];
            for (int w = 1; w < RED_WAVES; w++) t = t + red[k][w];
            partials[(size_t)blockIdx.x * NDIAG + k] = t;
        }
}

// Full sanity check (code/lb3d_sanity.F90:131-224) on the current
// populations.  Per block, in FP64: mass[3], mass flow[3], NaN / negative /
// too-fast / fluid site counts, the sums of the species velocities
// vel[c][a] = amass_c * sum_i g_i n_c,i c_i,a / rho_c (|rho_c| > 1e-12), and
// per (c, a) the largest |vel[c][a]| with its site.  Equal maxima keep the
// site that comes last in the reference's loop (x outermost, z innermost), as
// its `>=` update does: the key is that loop position, larger wins.
// Partials are reduced on the host.
template <typename R, bool Dev, int C>
__global__ __launch_bounds__(256) void k_sanity(Phys<R> P, const R *__restrict__ f,
                                                const uint8_t *__restrict__ rock,
                                                double *__restrict__ partials)
{
    __shared__ double red[NSAN][RED_WAVES];
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    double v[NSAN];
    for (int k = 0; k < NSAN; k++) v[k] = k < SAN_MAX ? 0.0 : -1.0;
    if (s < P.ns && !(rock && rock[s])) {  // fluid sites only (no_rock, code/lb3d_sanity.F90:134)
        int x, y, z;
        site_xyz(s, P.nx, P.ny, x, y, z);
        double key = ((double)x * (double)P.ny + (double)y) * (double)P.nz + (double)z;
        bool isnan_ = false, isneg = false, fast = false;
        v[SAN_FLUID] = 1.0;
        for (int c = 0; c < C; c++) {
            double rho = 0.0, m[3] = {0.0, 0.0, 0.0};
            for (int i = 0; i < 19; i++) {
                double n = (double)f[((uint32_t)c * 19u + (uint32_t)i) * P.ns + s];
                if (Dev) n = (i < 18 ? (double)P.rho0[c] / 36.0 : (double)P.rho0[c] / 3.0) + n;
                if (n != n) isnan_ = true;
                if (n < 0.0) isneg = true;
                double ng = n * (double)GD[i];
                rho = rho + ng;
                m[0] = m[0] + ng * (double)CX[i];
                m[1] = m[1] + ng * (double)CY[i];
                m[2] = m[2] + ng * (double)CZ[i];
            }
            v[c] = rho;
            double am = (double)P.amass[c];
            for (int a = 0; a < 3; a++) v[3 + a] = v[3 + a] + am * m[a];
            if (fabs(rho) > 1.0e-12) {
                double u[3];
                for (int a = 0; a < 3; a++) {
                    u[a] = m[a] * am / rho;
                    v[SAN_VSUM + 3 * c + a] = u[a];
                    double au = fabs(u[a]);
                    if (au >= 0.0) {  // false for NaN, which the reference's >= skips too
                        v[SAN_MAX + 3 * c + a] = au;
                        v[SAN_KEY + 3 * c + a] = key;
                    }
                }
                if (sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) > 0.1) fast = true;
            }
        }
        v[6] = isnan_ ? 1.0 : 0.0;
        v[7] = isneg ? 1.0 : 0.0;
        v[8] = fast ? 1.0 : 0.0;
    }
    const unsigned wv = threadIdx.x / 64u, ln = threadIdx.x % 64u;
    for (int k = 0; k < SAN_MAX; k++) {
        const double t = wave_sum(v[k]);
        if (ln == 0) red[k][wv] = t;
    }
    for (int k = 0; k < 9; k++) {
        double m = v[SAN_MAX + k], key = v[SAN_KEY + k];
        wave_argmax(m, key);
        if (ln == 0) {
            red[SAN_MAX + k][wv] = m;
            red[SAN_KEY + k][wv] = key;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < SAN_MAX; k++) {
            double t = red[k][0];
            for (int w = 1; w < RED_WAVES; w++) t = t + red[k][w];
            partials[(size_t)blockIdx.x * NSAN + k] = t;
        }
        for (int k = 0; k < 9; k++) {
            double m = red[SAN_MAX + k][0], key = red[SAN_KEY + k][0];
            for (int w = 1; w < RED_WAVES; w++) {
                const double m1 = red[SAN_MAX + k][w], k1 = red[SAN_KEY + k][w];
                if (m1 > m || (m1 == m && k1 > key)) {
                    m = m1;
                    key = k1;
                }
            }
            partials[(size_t)blockIdx.x * NSAN + SAN_MAX + k] = m;
            partials[(size_t)blockIdx.x * NSAN + SAN_KEY + k] = key;
        }
    }
}

template <typename R, bool Dev>
class LauncherT : public Launcher {
  public:
    explicit LauncherT(const lb3d_params &p)
        : P(make_phys<R>(p)), C(p.ncomp), le(p.boundary == LB3D_BC_LEES_EDWARDS)
    {
    }
    unsigned blocks() const { return (P.ns + BLOCK - 1) / BLOCK; }
    // grid of the stepping kernels: whole XCD rounds (xcd_tile order)
    unsigned tiles() const { return xcd_round(blocks()); }
    unsigned diag_blocks() const override { return blocks(); }
    size_t real_bytes() const override { return sizeof(R); }

    void moments(const DevBufs &b, hipStream_t s) override
    {
        const R *f = (const R *)b.f[b.cur];
        if (C == 3 && P.zh)
            k_moments<R, Dev, 3, false, false, true><<<tiles(), BLOCK, 0, s>>>(P, f, nullptr, (const R *)b.db,
                                                                               (R *)b.aux, xg, nullptr);
        else if (C == 3)
            k_moments<R, Dev, 3, false><<<tiles(), BLOCK, 0, s>>>(P, f, nullptr, (const R *)b.db, (R *)b.aux, xg,
                                                                  nullptr);
        else if (le)
            k_moments<R, Dev, 2, true><<<tiles(), BLOCK, 0, s>>>(P, f, (const R *)b.face_pc, nullptr, (R *)b.aux, xg,
                                                                 nullptr, (R *)b.edge_adv);
        else if (b.rock)
            k_moments<R, Dev, 2, false, true><<<tiles(), BLOCK, 0, s>>>(P, f, nullptr, nullptr, (R *)b.aux, xg,
                                                                        b.rock);
        else
            k_moments<R, Dev, 2, false><<<tiles(), BLOCK, 0, s>>>(P, f, nullptr, nullptr, (R *)b.aux, xg, nullptr);
    }
    void forces(const DevBufs &b, void *out, hipStream_t s) override
    {
        if (C == 3)
            k_forces<R, Dev, 3, false><<<blocks(), BLOCK, 0, s>>>(P, (const R *)b.aux, (const R *)b.db, nullptr,
                                                                  (R *)out);
        else if (b.rock)
            k_forces<R, Dev, 2, true><<<blocks(), BLOCK, 0, s>>>(P, (const R *)b.aux, nullptr, b.rock, (R *)out);
        else
            k_forces<R, Dev, 2, false><<<blocks(), BLOCK, 0, s>>>(P, (const R *)b.aux, nullptr, nullptr, (R *)out);
    }
    void dipole(const DevBufs &b, hipStream_t s) override
    {
        // the relaxed field goes to db_alt (the collide reads it there)
        if (C == 3) k_dipole<R, Dev><<<tiles(), BLOCK, 0, s>>>(P, (const R *)b.aux, (R *)b.db_alt, xg);
    }
    template <int BD, bool NT>
    void collide_bd(const DevBufs &b, hipStream_t s)
    {
        const R *f = (const R *)b.f[b.cur];
        R *fo = (R *)b.f[1 - b.cur];
        if (P.zh && C == 3)
            k_collide<R, Dev, 3, false, BD, false, true, NT><<<tiles(), BLOCK, 0, s>>>(
                P, f, nullptr, (const R *)b.aux, nullptr, (const R *)b.db_alt, fo, xg, nullptr);
        else if (P.zh && !le && !b.rock)
            k_collide<R, Dev, 2, false, BD, false, true, NT><<<tiles(), BLOCK, 0, s>>>(
                P, f, nullptr, (const R *)b.aux, nullptr, nullptr, fo, xg, nullptr);
        else if (C == 3)
            k_collide<R, Dev, 3, false, BD, false, false, NT><<<tiles(), BLOCK, 0, s>>>(
                P, f, nullptr, (const R *)b.aux, nullptr, (const R *)b.db_alt, fo, xg, nullptr);
        else if (le)
            k_collide<R, Dev, 2, true, BD, false, false, NT><<<tiles(), BLOCK, 0, s>>>(
                P, f, (const R *)b.face_pc, (const R *)b.aux, (const R *)b.face_aux, nullptr, fo, xg, nullptr,
                (R *)b.edge_pc, le_commit == LE_STATE_AS_IS ? nullptr : b.le, le_commit);
        else if (b.rock)
            k_collide<R, Dev, 2, false, BD, true><<<tiles(), BLOCK, 0, s>>>(P, f, nullptr, (const R *)b.aux,
                                                                            nullptr, nullptr, fo, xg, b.rock);
        else
            k_collide<R, Dev, 2, false, BD, false, false, NT><<<tiles(), BLOCK, 0, s>>>(
                P, f, nullptr, (const R *)b.aux, nullptr, nullptr, fo, xg, nullptr);
    }
    // the non-temporal form once one copy of the populations exceeds the
    // memory-side cache
    bool nt() const { return (uint64_t)C * 19u * P.ns * sizeof(R) > MALL_BYTES; }
    void collide(const DevBufs &b, hipStream_t s) override
    {
        const bool n = nt();
        if (P.bdist == 0)
            n ? collide_bd<0, true>(b, s) : collide_bd<0, false>(b, s);
        else if (P.bdist == 2)
            n ? collide_bd<2, true>(b, s) : collide_bd<2, false>(b, s);
        else
            n ? collide_bd<BD_GENERAL, true>(b, s) : collide_bd<BD_GENERAL, false>(b, s);
    }
    void le_faces(const DevBufs &b, bool post, bool compact, hipStream_t s, int adv) override
    {
        unsigned plane = (unsigned)P.ny * (unsigned)P.nz;
        unsigned nb = (plane + LE_SITES - 1) / LE_SITES;
        const R *f = (const R *)b.f[b.cur];
        R *fp = (R *)b.face_pc, *fa = (R *)b.face_aux;
        if (post && compact)
            k_le_faces<R, Dev, true, true><<<nb, 2 * LE_SITES, 0, s>>>(P, b.le, f, fp, fa, (const R *)b.edge_adv, adv);
        else if (post)
            k_le_faces<R, Dev, true><<<nb, 2 * LE_SITES, 0, s>>>(P, b.le, f, fp, fa, nullptr, adv);
        else if (compact)
            k_le_faces<R, Dev, false, true><<<nb, 2 * LE_SITES, 0, s>>>(P, b.le, f, fp, fa, (const R *)b.edge_pc, adv);
        else
            k_le_faces<R, Dev, false><<<nb, 2 * LE_SITES, 0, s>>>(P, b.le, f, fp, fa, nullptr, adv);
    }
    // the next two-pass collide stores le_advance(state, adv) unless adv is
    // LE_STATE_AS_IS (then it leaves the state alone)
    int le_commit = LE_STATE_AS_IS;
    void set_le_commit(int adv) override { le_commit = adv; }
    void le_gather_edges(const DevBufs &b, hipStream_t s) override
    {
        unsigned n = 2u * 2u * 19u * (unsigned)P.ny * (unsigned)P.nz;
        k_le_gather_edges<R><<<(n + BLOCK - 1) / BLOCK, BLOCK, 0, s>>>(P, (const R *)b.f[b.cur], (R *)b.edge_pc);
    }
    // LDS of k_fused2: the population stage plus the psi ring
    template <int TX, int TY>
    static constexpr bool fits()
    {
        return (38 * TX * TY + 8 * (TX + 2) * (TY + 2)) * sizeof(R) <= 160 * 1024;
    }
    // stage: 0 registers, 1 LDS, 2 reload (k_fused2's SM)
    template <int TX, int TY>
    bool fused_t(const DevBufs &b, int lz, int stage, hipStream_t s, bool launch)
    {
        if (stage == 1) {
            if constexpr (!fits<TX, TY>()) {
                return false;
            } else {
                if (launch) fused_go<TX, TY, 1>(b, lz, s);
                return true;
            }
        }
        if (stage == 0) {
            if (launch) fused_go<TX, TY, 0>(b, lz, s);
            return true;
        }
        if (stage == 2) {
            if (launch) fused_go<TX, TY, 2>(b, lz, s);
            return true;
        }
        return false;
    }
    template <int TX, int TY, int SM>
    void fused_go(const DevBufs &b, int lz, hipStream_t s)
    {
        unsigned grid = xcd_round(((P.nx + TX - 1) / TX) * ((P.ny + TY - 1) / TY) * ((P.nz + lz - 1) / lz));
        const R *f = (const R *)b.f[b.cur];
        R *fo = (R *)b.f[1 - b.cur];
        if (le)
            k_fused2<R, Dev, true, TX, TY, SM><<<grid, TX * TY, 0, s>>>(P, f, (const R *)b.face_pc,
                                                                         (const R *)b.face_aux, fo, lz, xg);
        else
            k_fused2<R, Dev, false, TX, TY, SM><<<grid, TX * TY, 0, s>>>(P, f, nullptr, nullptr, fo, lz, xg);
    }
    // LDS of k_fused3: its rings of moments, psi and dipoles
    template <int TX, int TY>
    static constexpr bool fits3()
    {
        return Fused3Dims<TX, TY>::template lds_bytes<R>() <= 160 * 1024;
    }
    template <int TX, int TY>
    bool fused3_t(const DevBufs &b, int lz, hipStream_t s, bool launch)
    {
        if constexpr (!fits3<TX, TY>()) {
            return false;
        } else {
            if (launch) {
                unsigned grid =
                    xcd_round(((P.nx + TX - 1) / TX) * ((P.ny + TY - 1) / TY) * ((P.nz + lz - 1) / lz));
                const R *f = (const R *)b.f[b.cur];
                R *fo = (R *)b.f[1 - b.cur];
                if (P.bdist == 0)
                    k_fused3<R, Dev, TX, TY, 0><<<grid, TX * TY, 0, s>>>(P, f, (const R *)b.db, (R *)b.db_alt, fo,
                                                                          lz, xg);
                else
                    k_fused3<R, Dev, TX, TY, 2><<<grid, TX * TY, 0, s>>>(P, f, (const R *)b.db, (R *)b.db_alt, fo,
                                                                          lz, xg);
            }
            return true;
        }
    }
    bool fused_dispatch(const DevBufs &b, int tx, int ty, int lz, int st, hipStream_t s, bool launch)
    {
        if (walls || P.zh) return false;  // walls and the Zou-He boundary run on the two-pass kernels
        if (C == 3) {
            // ternary: the reload form only, periodic lattices, nz >= 3 for the +-2 plane wrap
            if (le || lz < 1 || st != 2 || P.nz < 3 || P.nx < 3 || P.ny < 3) return false;
            // k_fused3 addresses the populations with 32-bit byte offsets
            if ((uint64_t)3 * 19 * P.ns * sizeof(R) > 0xffffffffull) return false;
            if (P.bdist != 0 && P.bdist != 2) return false;
            if (tx == 32 && ty == 4) return fused3_t<32, 4>(b, lz, s, launch);
            if (tx == 32 && ty == 8) return fused3_t<32, 8>(b, lz, s, launch);
            if (tx == 16 && ty == 16) return fused3_t<16, 16>(b, lz, s, launch);
            if (tx == 64 && ty == 4) return fused3_t<64, 4>(b, lz, s, launch);
            if (tx == 64 && ty == 8) return fused3_t<64, 8>(b, lz, s, launch);
            if (tx == 32 && ty == 16) return fused3_t<32, 16>(b, lz, s, launch);
            return false;
        }
        if (C != 2 || lz < 1 || P.bdist != 2) return false;
        if (tx == 64 && ty == 4) return fused_t<64, 4>(b, lz, st, s, launch);
        if (tx == 32 && ty == 8) return fused_t<32, 8>(b, lz, st, s, launch);
        if (tx == 64 && ty == 8) return fused_t<64, 8>(b, lz, st, s, launch);
        if (tx == 32 && ty == 16) return fused_t<32, 16>(b, lz, st, s, launch);
        if (tx == 128 && ty == 4) return fused_t<128, 4>(b, lz, st, s, launch);
        if (tx == 32 && ty == 4) return fused_t<32, 4>(b, lz, st, s, launch);
        if (tx == 128 && ty == 2) return fused_t<128, 2>(b, lz, st, s, launch);
        if (tx == 256 && ty == 2) return fused_t<256, 2>(b, lz, st, s, launch);
        return false;
    }
    bool fused(const DevBufs &b, int tx, int ty, int lz, int stage, hipStream_t s) override
    {
        return fused_dispatch(b, tx, ty, lz, stage, s, true);
    }
    bool fused_supported(int tx, int ty, int stage) override
    {
        DevBufs none;
        return fused_dispatch(none, tx, ty, 1, stage, nullptr, false);
    }
    void init_noise(const DevBufs &b, uint64_t seed, const double m[3], const double a[3], int dip,
                    hipStream_t s) override
    {
        R *f = (R *)b.f[b.cur];
        double r0[3] = {(double)P.rho0[0], (double)P.rho0[1], (double)P.rho0[2]};
        if (C == 3)
            k_init_noise<R, Dev, 3><<<blocks(), BLOCK, 0, s>>>(P, seed, m[0], m[1], m[2], a[0], a[1], a[2], r0[0],
                                                                r0[1], r0[2], dip, f, (R *)b.db);
        else
            k_init_noise<R, Dev, 2><<<blocks(), BLOCK, 0, s>>>(P, seed, m[0], m[1], m[2], a[0], a[1], a[2], r0[0],
                                                                r0[1], r0[2], dip, f, nullptr);
    }
    void sanity(const DevBufs &b, double *partials, hipStream_t s) override
    {
        const R *f = (const R *)b.f[b.cur];
        if (C == 3)
            k_sanity<R, Dev, 3><<<blocks(), BLOCK, 0, s>>>(P, f, b.rock, partials);
        else
            k_sanity<R, Dev, 2><<<blocks(), BLOCK, 0, s>>>(P, f, b.rock, partials);
    }
    void diagnostics(const DevBufs &b, double *partials, hipStream_t s) override
    {
        const R *f = (const R *)b.f[b.cur];
        if (C == 3)
            k_diagnostics<R, Dev, 3><<<blocks(), BLOCK, 0, s>>>(P, f, b.rock, partials);
        else
            k_diagnostics<R, Dev, 2><<<blocks(), BLOCK, 0, s>>>(P, f, b.rock, partials);
    }

  private:
    Phys<R> P;
    int C;
    bool le;
};

template <typename R, bool Dev>
class BatchLauncherT : public BatchLauncher {
  public:
    int bd_mask = 0;  // equilibrium forms present in the batch (bit b: bdist b)
    bool nt = false;  // non-temporal population stores: the populations exceed the memory-side cache
    void set_bdist_mask(int m) override { bd_mask = m; }
    void set_pop_bytes(uint64_t bytes) override { nt = bytes > MALL_BYTES; }
    size_t lat_bytes() const override { return sizeof(BatchLat<R>); }
    void fill_lat(void *dst, const lb3d_params &p, uint64_t f_off, uint64_t aux_off,
                  uint64_t db_off) const override
    {
        BatchLat<R> L{};
        L.P = make_phys<R>(p);
        L.f_off = f_off;
        L.aux_off = aux_off;
        L.db_off = db_off;
        std::memcpy(dst, &L, sizeof(L));
    }
    void phase(int ph, const void *lat, const void *items, uint32_t nitems, unsigned nwg, const void *f,
               void *fout, void *aux, void *db, int ncomp, hipStream_t s, bool packed) override
    {
        const BatchLat<R> *L = (const BatchLat<R> *)lat;
        const uint2 *it = (const uint2 *)items;
        if (ph == 0) {
            if (ncomp == 3 && packed)
                k_batch_moments<R, Dev, 3, true>
                    <<<nwg, BLOCK, 0, s>>>(L, it, nitems, (const R *)f, (const R *)db, (R *)aux, xg);
            else if (ncomp == 3)
                k_batch_moments<R, Dev, 3><<<nwg, BLOCK, 0, s>>>(L, it, nitems, (const R *)f, (const R *)db, (R *)aux, xg);
            else
                k_batch_moments<R, Dev, 2><<<nwg, BLOCK, 0, s>>>(L, it, nitems, (const R *)f, nullptr, (R *)aux, xg);
        } else if (ph == 1) {
            if (ncomp == 3 && packed)
                k_batch_dipole<R, Dev, true><<<nwg, BLOCK, 0, s>>>(L, it, nitems, (const R *)aux, (R *)db, xg);
            else if (ncomp == 3)
                k_batch_dipole<R, Dev><<<nwg, BLOCK, 0, s>>>(L, it, nitems, (const R *)aux, (R *)db, xg);
        } else {
            // the ternary collide reads the packed layout only: the moments
            // phase before it is launched with packed = true
            // every lattice's collide runs in the launch of its own equilibrium
            // form: bdist 0, bdist 2, or one launch for all the others
            for (int bd : {0, 2, BD_GENERAL}) {
                if (!(bd == BD_GENERAL ? (bd_mask & ~0b101) : (bd_mask & (1 << bd)))) continue;
                if (nt)
                    collide_bd<true>(bd, L, it, nitems, nwg, f, fout, aux, db, ncomp, s);
                else
                    collide_bd<false>(bd, L, it, nitems, nwg, f, fout, aux, db, ncomp, s);
            }
        }
    }
    // FP32 ternary batches whose lattices all take bdist 0 or 2
    bool cm_supported(int ncomp) const override { return Dev && ncomp == 3 && !(bd_mask & ~0b101); }
    void cm_collide(const void *lat, const void *items, uint32_t nitems, unsigned nwg, int threads, const void *f,
                    void *fout, const void *aux_in, void *aux_out, const void *db, hipStream_t s) override
    {
        if constexpr (Dev) {
            const BatchLat<R> *L = (const BatchLat<R> *)lat;
            const uint4 *it = (const uint4 *)items;
            // one launch per equilibrium form present; a launch skips the
            // tiles of the other form
            auto go = [&](auto bd, auto nthr) {
                constexpr int T = decltype(nthr)::value;
                k_batch_cm<R, Dev, decltype(bd)::value, T><<<nwg, T, 0, s>>>(
                    L, it, nitems, (const R *)f, (const R *)aux_in, (const R *)db, (R *)fout, (R *)aux_out, xg);
            };
            using T256 = std::integral_constant<int, 256>;
            using T512 = std::integral_constant<int, 512>;
            for (int bd : {0, 2}) {
                if (!(bd_mask & (1 << bd))) continue;
                if (bd == 0)
                    threads == 512 ? go(std::integral_constant<int, 0>{}, T512{})
                                   : go(std::integral_constant<int, 0>{}, T256{});
                else
                    threads == 512 ? go(std::integral_constant<int, 2>{}, T512{})
                                   : go(std::integral_constant<int, 2>{}, T256{});
            }
        }
    }
    void edge_moments(const void *lat, const void *items, uint32_t nitems, unsigned nwg, const void *sites,
                      const void *f, const void *db, void *aux, hipStream_t s) override
    {
        if constexpr (Dev)
            k_batch_edge_moments<R, Dev><<<nwg, BLOCK, 0, s>>>((const BatchLat<R> *)lat, (const uint2 *)items,
                                                               nitems, (const uint32_t *)sites, (const R *)f,
                                                               (const R *)db, (R *)aux, xg);
    }
    template <bool NT>
    void collide_bd(int bd, const BatchLat<R> *L, const uint2 *it, uint32_t nitems, unsigned nwg, const void *f,
                    void *fout, void *aux, void *db, int ncomp, hipStream_t s)
    {
        if (ncomp == 3 && bd == 0)
            k_batch_collide<R, Dev, 3, 0, NT, true><<<nwg, BLOCK, 0, s>>>(L, it, nitems, (const R *)f, (const R *)aux,
                                                                           (const R *)db, (R *)fout, xg);
        else if (ncomp == 3 && bd == BD_GENERAL)
            k_batch_collide<R, Dev, 3, BD_GENERAL, NT, true><<<nwg, BLOCK, 0, s>>>(
                L, it, nitems, (const R *)f, (const R *)aux, (const R *)db, (R *)fout, xg);
        else if (bd == BD_GENERAL)
            k_batch_collide<R, Dev, 2, BD_GENERAL, NT><<<nwg, BLOCK, 0, s>>>(L, it, nitems, (const R *)f,
                                                                              (const R *)aux, nullptr, (R *)fout, xg);
        else if (ncomp == 3)
            k_batch_collide<R, Dev, 3, 2, NT, true><<<nwg, BLOCK, 0, s>>>(L, it, nitems, (const R *)f, (const R *)aux,
                                                                           (const R *)db, (R *)fout, xg);
        else if (bd == 0)
            k_batch_collide<R, Dev, 2, 0, NT><<<nwg, BLOCK, 0, s>>>(L, it, nitems, (const R *)f, (const R *)aux,
                                                                     nullptr, (R *)fout, xg);
        else
            k_batch_collide<R, Dev, 2, 2, NT><<<nwg, BLOCK, 0, s>>>(L, it, nitems, (const R *)f, (const R *)aux,
                                                                     nullptr, (R *)fout, xg);
    }
};

}  // namespace lb3d
