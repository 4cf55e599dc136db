// lb3d_kernels.h -- the per-timestep kernels, templated on the storage form.
//
// Storage (structure of arrays, x fastest, no halo; periodic boundaries are
// index wraps):
//   f[(c*19 + i)*ns + s]   populations, two buffers (pull: read one, write other)
//   aux[k*ns + s]          moments of the streamed state, written by k_moments:
//                          binary  k = psi_r, psi_b, rho_r, rho_b
//                          ternary k = psi_r, psi_b, psi_s, rho_r, rho_b, rho_s, da_x, da_y, da_z
//   db[k*ns + s]           ternary dipole field
//   face_pc[((f*C + c)*19 + i)*ny*nz + z*ny + y]  Lees-Edwards x-halo planes
//                          (f = 0: x=0 plane, f = 1: x=nx+1 plane) of the
//                          post-collision populations, built by k_le_faces<false>
//   face_aux[(f*4 + k)*ny*nz + z*ny + y]  moments of the post-advection
//                          x-halo planes (binary), built by k_le_faces<true>
//
// Form FP64 (Dev = false, R = double): the reference's values and arithmetic
// order; compile with -ffp-contract=off.  Form FP32_DEV (Dev = true,
// R = float): f holds n_i - w_i*rho0_c, rho entries hold rho - rho0_c, psi
// entries hold psi(rho) - psi(rho0_c).
#pragma once

#include "lb3d_lattice.h"

namespace lb3d {

// collide kernels are compiled per equilibrium form: bdist 0 and 2 inline,
// every other form (1, 3, 4, 5, 6) through boltz_general at run time
constexpr int BD_GENERAL = 9;

// Moments of the streamed state (the aux field).  Binary: planes psi_r,
// psi_b, rho_r, rho_b.  Ternary: planes psi_r, psi_b, psi_s, rho_r, rho_b,
// rho_s, then at a 4-element boundary one 4-vector per site of what the
// dipole relaxation reads at its 18 neighbours, (rho_b - rho_r, da_xyz).
// The batch's three-pass plan writes its ternary field packed: in place of
// the six planes, one 4-vector (psi_r, psi_b, psi_s, rho_s) per site from the
// start of the field (8 ns elements <= aux_elems), then the same 4-vectors.
constexpr int NAUX2 = 4;
template <typename R>
struct alignas(4 * sizeof(R)) V4 {
    R a, b, c, d;
};
__host__ __device__ constexpr uint64_t aux3_vec_off(uint64_t ns) { return (6 * ns + 3) / 4 * 4; }

// Summation order of the FP32 ternary moments: the directions grouped by
// their z component, c_z = +1 first, then 0, then -1, ascending within a
// group.  A kernel that gathers the moments of the next step from the
// post-collision populations of one z-plane at a time (k_batch_cm) receives
// them in this order, so every FP32 ternary kernel sums them in it and all
// plans round alike.  FP64 and binary lattices keep the reference's order.
__device__ constexpr int ZORD[19] = {4, 8, 12, 14, 16, 0, 1, 2, 3, 6, 7, 10, 11, 18, 5, 9, 13, 15, 17};
template <bool ZG>
__device__ __forceinline__ constexpr int dir_at(int ii)
{
    return ZG ? ZORD[ii] : ii;
}
__host__ __device__ constexpr uint64_t aux_elems(int ncomp, uint64_t ns)
{
    return ncomp == 3 ? aux3_vec_off(ns) + 4 * ns : (uint64_t)NAUX2 * ns;
}

// Work-groups are dispatched round-robin over the chip's 8 XCDs (XCD =
// work-group index mod 8), each XCD with its own L2.  xcd_tile maps the
// work-group index b of a grid of n8 (a multiple of 8) to a tile so that each
// XCD runs chunks of xg consecutive tiles: the neighbour rows and planes that
// the pulls and stencils of a tile re-read are then fetched by the same L2.
// xg = 1 is the plain order; xg = 0 (or >= n8/8) gives each XCD one
// contiguous run.  A bijection on [0, n8); tiles >= the tile count idle.
constexpr uint32_t NXCD = 8;
__host__ __device__ constexpr uint32_t xcd_round(uint32_t n) { return (n + NXCD - 1) / NXCD * NXCD; }
__host__ __device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n8, uint32_t xg)
{
    uint32_t per = n8 / NXCD, j = b / NXCD, k = b % NXCD;
    if (xg == 0 || xg > per) xg = per;
    uint32_t c = j / xg, o = j - c * xg;
    uint32_t w = per - c * xg < xg ? per - c * xg : xg;  // the last chunk may be narrower
    return c * NXCD * xg + k * w + o;
}

__device__ __forceinline__ void site_xyz(uint32_t s, int nx, int ny, int &x, int &y, int &z)
{
    uint32_t t = s / (uint32_t)nx;
    x = (int)(s - t * (uint32_t)nx);
    z = (int)(t / (uint32_t)ny);
    y = (int)(t - (uint32_t)z * (uint32_t)ny);
}

__device__ __forceinline__ uint32_t sidx(int x, int y, int z, int nx, int ny)
{
    return (uint32_t)x + (uint32_t)nx * ((uint32_t)y + (uint32_t)ny * (uint32_t)z);
}

// The Lees-Edwards state of the next step (le_neighbours with the stage-6
// shear update, code/lb3d_bc_leesedwards.F90:86-140, code/lb3d_bc.F90:110):
// nt_set >= 0 sets the step counter, -1 advances it by one.  FP64.
__device__ __forceinline__ LEState le_advance(LEState s, int nz, int nt_set)
{
    s.nt = nt_set >= 0 ? nt_set : s.nt + 1;
    // (cos(0) = 1 exactly: a steady shear skips the cosine)
    s.shear_u = s.shear_omega == 0.0 ? s.shear_tmp : s.shear_tmp * cos(s.shear_omega * (double)s.nt);
    s.shear_sum = s.shear_sum + s.shear_u;
    // lelo = mod(int(2.*shear_sum), nz): truncation, sign of the dividend
    long t = (long)(2.0 * s.shear_sum);
    s.lelo = (int)(t % (long)nz);
    double lf = 2.0 * s.shear_sum;
    s.lefrac = fabs(lf - trunc(lf));
    return s;
}

// Loads and stores through a wave-uniform plane base plus a 32-bit byte offset
// per lane: the plane bases live in SGPRs and each access needs one VGPR for
// its offset (the global "saddr" addressing form) instead of a 64-bit address.
template <typename R>
__device__ __forceinline__ R ld_plane(const R *__restrict__ plane, uint32_t site)
{
    return *(const R *)((const char *)plane + site * (uint32_t)sizeof(R));
}
template <typename R>
__device__ __forceinline__ void st_plane(R *__restrict__ plane, uint32_t site, R v)
{
    *(R *)((char *)plane + site * (uint32_t)sizeof(R)) = v;
}
// Population loads and stores with an optional non-temporal hint (NT): the
// streamed populations are touched once per pass, and the hint keeps them
// from displacing the stencil fields that neighbouring sites re-read from L2.
// It pays only when the populations exceed the memory-side cache (MI355X:
// 256 MiB of Infinity Cache): a smaller lattice written without the hint is
// read back from that cache by the next step.
constexpr uint64_t MALL_BYTES = 256ull << 20;
template <bool NT, typename R>
__device__ __forceinline__ R ld_pop(const R *__restrict__ plane, uint32_t site)
{
    if constexpr (NT)
        return __builtin_nontemporal_load((const R *)((const char *)plane + site * (uint32_t)sizeof(R)));
    else
        return ld_plane(plane, site);
}
template <bool NT, typename R>
__device__ __forceinline__ void st_pop(R *__restrict__ plane, uint32_t site, R v)
{
    if constexpr (NT)
        __builtin_nontemporal_store(v, (R *)((char *)plane + site * (uint32_t)sizeof(R)));
    else
        st_plane(plane, site, v);
}
template <typename R>
__device__ __forceinline__ const R *plane_of(const R *__restrict__ base, uint32_t k, uint32_t n)
{
    return base + (size_t)k * (size_t)n;
}

// Population i of component c pulled into site (x,y,z): the value at
// (x-c_i, y-c_i, z-c_i).  With Lees-Edwards a source outside 0..nx-1 in x
// reads the x-halo plane face_pc.
template <typename R, int C, bool LE, bool NT = false>
__device__ __forceinline__ R pull(const R *__restrict__ f, const R *__restrict__ face_pc,
                                  const Phys<R> &P, int c, int i, int x, int y, int z)
{
    int xs = x - CX[i], ys = wrap1(y - CY[i], P.ny), zs = wrap1(z - CZ[i], P.nz);
    if (LE) {
        if (xs < 0 || xs >= P.nx) {
            int fc = xs < 0 ? 0 : 1;
            uint32_t plane = (uint32_t)P.ny * (uint32_t)P.nz;
            return ld_plane(plane_of(face_pc, (uint32_t)(fc * C + c) * 19u + (uint32_t)i, plane),
                            (uint32_t)zs * (uint32_t)P.ny + (uint32_t)ys);
        }
    } else {
        xs = wrap1(xs, P.nx);
    }
    return ld_pop<NT>(plane_of(f, (uint32_t)c * 19u + (uint32_t)i, P.ns), sidx(xs, ys, zs, P.nx, P.ny));
}

// Pull through one base pointer: element (c*19 + i)*ns + site, addressed as a
// 32-bit byte offset.  The plane offset is a scalar product instead of a
// 64-bit base per population plane, which a kernel holding the moments of
// several planes at once could not keep in scalar registers.  Valid while the
// populations take less than 4 GiB.
template <typename R, int C>
__device__ __forceinline__ R pull_flat(const R *__restrict__ f, const Phys<R> &P, int c, int i, int x, int y, int z)
{
    int xs = wrap1(x - CX[i], P.nx), ys = wrap1(y - CY[i], P.ny), zs = wrap1(z - CZ[i], P.nz);
    return ld_plane(f, ((uint32_t)c * 19u + (uint32_t)i) * P.ns + sidx(xs, ys, zs, P.nx, P.ny));
}

// Pull with walls (periodic binary lattices with rock sites): a rock site
// keeps its own populations (code/lb3d_advection.F90:126), and a fluid site
// whose source x - c_i is rock takes its own n_bounce(i) instead, the
// half-way bounce-back of code/lb3d_bc_bounce.F90:53-90.
template <typename R, int C>
__device__ __forceinline__ R pull_rock(const R *__restrict__ f, const uint8_t *__restrict__ rock,
                                       const Phys<R> &P, int c, int i, int x, int y, int z, bool self_rock)
{
    const uint32_t s = sidx(x, y, z, P.nx, P.ny);
    if (self_rock) return ld_plane(plane_of(f, (uint32_t)c * 19u + (uint32_t)i, P.ns), s);
    int xs = wrap1(x - CX[i], P.nx), ys = wrap1(y - CY[i], P.ny), zs = wrap1(z - CZ[i], P.nz);
    uint32_t q = sidx(xs, ys, zs, P.nx, P.ny);
    if (rock[q]) return ld_plane(plane_of(f, (uint32_t)c * 19u + (uint32_t)BOUNCE[i], P.ns), s);
    return ld_plane(plane_of(f, (uint32_t)c * 19u + (uint32_t)i, P.ns), q);
}

// Value k of the aux field at the neighbour (x+dx, y+dy, z+dz); for LE the
// x-halo neighbours read face_aux.
template <typename R, bool LE>
__device__ __forceinline__ R aux_at(const R *__restrict__ aux, const R *__restrict__ face_aux,
                                    const Phys<R> &P, int k, int x, int y, int z)
{
    int ys = wrap1(y, P.ny), zs = wrap1(z, P.nz);
    if (LE) {
        if (x < 0 || x >= P.nx) {
            int fc = x < 0 ? 0 : 1;
            uint32_t plane = (uint32_t)P.ny * (uint32_t)P.nz;
            return ld_plane(plane_of(face_aux, (uint32_t)(fc * 4 + k), plane),
                            (uint32_t)zs * (uint32_t)P.ny + (uint32_t)ys);
        }
    } else {
        x = wrap1(x, P.nx);
    }
    return ld_plane(plane_of(aux, (uint32_t)k, P.ns), sidx(x, ys, zs, P.nx, P.ny));
}

// psi(rho) = 1 - exp(-rho) (get_psi, code/lb3d_force_shanchen.F90:334-353);
// deviation form: psi(rho0+d) - psi(rho0) = -exp(-rho0)*expm1(-d).
template <bool Dev, typename R>
__device__ __forceinline__ R psi_of(R rho, R emr0)
{
    if constexpr (Dev)
        return -emr0 * expm1f(-rho);
    else
        return R(1.0) - exp(-rho);
}

// --------------------------------------------------------------------------
// k_moments: streaming (pull) and the moments of the streamed populations,
// code/lb3d_advection.F90:54-151 with lb3d_update_density
// (code/lb3d_analysis.F90:64-76); ternary also the advected dipole
// da = sum_i g_i n_s,i(x-c_i) db(x-c_i) (:130-148).  The streamed
// populations themselves are not stored: k_collide pulls them again.
// One term of the advected dipole: the streamed surfactant population v of
// direction i (stored form) times the dipole field at its source x - c_i.
// da += ng * db: the FP32 form with an explicit fused multiply-add, so that
// every kernel rounds it alike whatever the compiler contracts; FP64 as the
// reference writes it (separate multiply and add).
template <bool Dev, typename R>
__device__ __forceinline__ void da_term(R ng, R b0, R b1, R b2, R &da0, R &da1, R &da2)
{
    if constexpr (Dev) {
        da0 = __builtin_fmaf(ng, b0, da0);
        da1 = __builtin_fmaf(ng, b1, da1);
        da2 = __builtin_fmaf(ng, b2, da2);
    } else {
        da0 = da0 + ng * b0;
        da1 = da1 + ng * b1;
        da2 = da2 + ng * b2;
    }
}

template <typename R, bool Dev>
__device__ __forceinline__ void acc_da(const Phys<R> &P, const R *__restrict__ db, int i, R v, int x, int y, int z,
                                       R &da0, R &da1, R &da2)
{
    R ns_ = Dev ? (i < 18 ? P.wrho0_move[2] : P.wrho0_rest[2]) + v : v;
    R ng = ns_ * R(GD[i]);
    int xs = wrap1(x - CX[i], P.nx), ys = wrap1(y - CY[i], P.ny), zs = wrap1(z - CZ[i], P.nz);
    uint32_t q = sidx(xs, ys, zs, P.nx, P.ny);
    da_term<Dev>(ng, ld_plane(db, q), ld_plane(plane_of(db, 1u, P.ns), q), ld_plane(plane_of(db, 2u, P.ns), q),
                 da0, da1, da2);
}

// Zou-He flux rule, lbe_invade_zouandheflux (code/lb3d_invasion.F90:2143-2457)
// with ux = uy = 0, on the streamed populations n (stored form) of one
// component at a site of the z = 0 plane (low) or the z = nz-1 plane: the
// populations entering the lattice are rebuilt so that the z flux is uz
// (itype < 0: uz is a mass flow, divided by the boundary density).  F = g n
// in the reference's 1-based direction numbering.  In the deviation form
// F = g (n - w rho0): the rest state satisfies every relation, so they hold
// for the deviations as written, with the boundary density rho0 + their sum.
#define ZF(k) F[(k) - 1]
template <typename R, bool Dev>
__device__ __forceinline__ void zou_he_site(R (&n)[19], R uz, int itype, bool low, R rho0)
{
    R F[19];
#pragma unroll
    for (int s = 0; s < 19; s++) F[s] = n[s] * R(GD[s]);
    const R ux = R(0), uy = R(0);
    R rho;
    if (low) {  // :2179-2214
        R sum = ZF(19) + ZF(1) + ZF(2) + ZF(3) + ZF(4) + ZF(7) + ZF(11) + ZF(12) + ZF(8) +
                R(2.0) * (ZF(6) + ZF(10) + ZF(14) + ZF(16) + ZF(18));
        if constexpr (Dev) sum = rho0 + sum;
        if (itype < 0) {
            rho = sum + uz;
            uz = rho > R(0) ? uz / rho : R(0);
        } else {
            rho = sum / (R(1.0) - uz);
        }
        ZF(5) = ZF(6) + uz * rho / R(3.0);
        if (itype == 0) {
            ZF(15) = ZF(18) - (ZF(3) - ZF(4)) / R(4.0) + rho * uz / R(6.0);
            ZF(17) = ZF(16) + (ZF(3) - ZF(4)) / R(4.0) + rho * uz / R(6.0);
            ZF(9) = ZF(14) - (ZF(1) - ZF(2)) / R(4.0) + rho * uz / R(6.0);
            ZF(13) = ZF(10) + (ZF(1) - ZF(2)) / R(4.0) + rho * uz / R(6.0);
        } else {
            ZF(15) = ZF(18) - (ZF(7) + ZF(11) + ZF(3) - ZF(8) - ZF(12) - ZF(4)) / R(2.0) + uz * rho / R(6.0) +
                     uy * rho / R(2.0);
            ZF(17) = ZF(16) - (ZF(8) + ZF(12) + ZF(4) - ZF(7) - ZF(11) - ZF(3)) / R(2.0) + uz * rho / R(6.0) -
                     uy * rho / R(2.0);
            ZF(9) = ZF(14) - (ZF(7) + ZF(8) + ZF(1) - ZF(11) - ZF(12) - ZF(2)) / R(2.0) + uz * rho / R(6.0) +
                    ux * rho / R(2.0);
            ZF(13) = ZF(10) - (ZF(11) + ZF(12) + ZF(2) - ZF(7) - ZF(8) - ZF(1)) / R(2.0) + uz * rho / R(6.0) -
                     ux * rho / R(2.0);
        }
    } else {  // :2318-2352
        R sum = ZF(19) + ZF(1) + ZF(2) + ZF(3) + ZF(4) + ZF(7) + ZF(11) + ZF(12) + ZF(8) +
                R(2.0) * (ZF(5) + ZF(9) + ZF(13) + ZF(15) + ZF(17));
        if constexpr (Dev) sum = rho0 + sum;
        if (itype < 0) {
            rho = sum - uz;
            uz = rho > R(0) ? uz / rho : R(0);
        } else {
            rho = sum / (uz + R(1.0));
        }
        ZF(6) = ZF(5) - uz * rho / R(3.0);
        if (itype == 0) {
            ZF(18) = ZF(15) + (ZF(3) - ZF(4)) / R(4.0) - rho * uz / R(6.0);
            ZF(16) = ZF(17) - (ZF(3) - ZF(4)) / R(4.0) - rho * uz / R(6.0);
            ZF(14) = ZF(9) + (ZF(1) - ZF(2)) / R(4.0) - rho * uz / R(6.0);
            ZF(10) = ZF(13) - (ZF(1) - ZF(2)) / R(4.0) - rho * uz / R(6.0);
        } else {
            ZF(18) = ZF(15) + (ZF(7) + ZF(11) + ZF(3) - ZF(8) - ZF(12) - ZF(4)) / R(2.0) - uz * rho / R(6.0) -
                     uy * rho / R(2.0);
            ZF(16) = ZF(17) + (ZF(8) + ZF(12) + ZF(4) - ZF(7) - ZF(11) - ZF(3)) / R(2.0) - uz * rho / R(6.0) +
                     uy * rho / R(2.0);
            ZF(14) = ZF(9) + (ZF(7) + ZF(8) + ZF(1) - ZF(11) - ZF(12) - ZF(2)) / R(2.0) - uz * rho / R(6.0) -
                     ux * rho / R(2.0);
            ZF(10) = ZF(13) + (ZF(11) + ZF(12) + ZF(2) - ZF(7) - ZF(8) - ZF(1)) / R(2.0) - uz * rho / R(6.0) +
                     ux * rho / R(2.0);
        }
    }
#pragma unroll
    for (int s = 0; s < 19; s++) n[s] = F[s] / R(GD[s]);
}
#undef ZF

// The rule on both planes a site may sit on (nz = 1: the low one first).
template <typename R, bool Dev>
__device__ __forceinline__ void zou_he_planes(const Phys<R> &P, R (&n)[19], int c, int z)
{
    if (z == 0) zou_he_site<R, Dev>(n, P.zh_flux[c], P.zh_type, true, P.rho0[c]);
    if (z == P.nz - 1) zou_he_site<R, Dev>(n, P.zh_flux[c], P.zh_type, false, P.rho0[c]);
}

// Lees-Edwards edge columns: the populations of the x = nx-1 (col 0) and
// x = 0 (col 1) sites in a compact site-major layout
// edge[(col*ny*nz + z*ny + y)*38 + c*19 + i].  A site's 38 values share two
// or three cache lines, so the one lane of a row that writes them touches
// those lines only, and the halo-plane kernels read them contiguously (in
// the lattice layout each of them lies on its own cache line).
template <typename T, typename R>
__device__ __forceinline__ T *edge_site(T *edge, const Phys<R> &P, int col, uint32_t q)
{
    return edge + ((uint32_t)col * ((uint32_t)P.ny * (uint32_t)P.nz) + q) * 38u;
}

// Stores population (c, i) of an edge-column site into the compact layout
// from the register that already holds it; e is null off the edge columns
// (and when no edge buffer is given), so the other lanes store nothing.
template <typename R>
struct EdgeSink {
    R *e;
    __device__ __forceinline__ void operator()(int c, int i, R v) const
    {
        if (e) e[c * 19 + i] = v;
    }
};
template <typename R>
__device__ __forceinline__ EdgeSink<R> edge_sink(R *edge, const Phys<R> &P, int x, int y, int z)
{
    R *e = nullptr;
    if (edge && (x == 0 || x == P.nx - 1))  // nx >= 2: a site is at most one column
        e = edge_site(edge, P, x == 0 ? 1 : 0, (uint32_t)z * (uint32_t)P.ny + (uint32_t)y);
    return EdgeSink<R>{e};
}

template <typename R, bool Dev, int C, bool LE, bool Rock = false, bool ZH = false, bool Flat = false>
__device__ __forceinline__ void moments_core(const Phys<R> &P, const R *__restrict__ f,
                                             const R *__restrict__ face_pc,
                                             const R *__restrict__ db, int x, int y, int z, R (&rho)[C],
                                             R (&da)[3], const uint8_t *__restrict__ rock = nullptr,
                                             R *__restrict__ edge = nullptr)
{
    static_assert(!Rock || (C == 2 && !LE), "walls: periodic binary lattices");
#pragma unroll
    for (int c = 0; c < C; c++) rho[c] = R(0);
    R da0 = R(0), da1 = R(0), da2 = R(0);
    if constexpr (ZH && C == 3) {
        if (z == 0 || z == P.nz - 1) {
            // Zou-He planes: rho of the streamed populations, the advected
            // dipole from the rebuilt surfactant populations
            // (code/lb3d_invasion.F90:2216-2228)
            R ns[19];
#pragma unroll
            for (int i = 0; i < 19; i++)
#pragma unroll
                for (int c = 0; c < C; c++) {
                    R v = pull<R, C, LE>(f, face_pc, P, c, i, x, y, z);
                    rho[c] = rho[c] + v * R(GD[i]);
                    if (c == 2) ns[i] = v;
                }
            zou_he_planes<R, Dev>(P, ns, 2, z);
#pragma unroll
            for (int i = 0; i < 19; i++) acc_da<R, Dev>(P, db, i, ns[i], x, y, z, da0, da1, da2);
            da[0] = da0;
            da[1] = da1;
            da[2] = da2;
            return;
        }
    }
    if constexpr (LE && C == 2) {
        if (edge) {
            // Lees-Edwards binary, with the streamed edge-column populations
            // kept for the post-advection halo planes: every pull is issued
            // before the first (edge-lane) store, so the stores wait once
            R n[C][19];
#pragma unroll
            for (int i = 0; i < 19; i++)
#pragma unroll
                for (int c = 0; c < C; c++) n[c][i] = pull<R, C, LE>(f, face_pc, P, c, i, x, y, z);
#pragma unroll
            for (int i = 0; i < 19; i++)
#pragma unroll
                for (int c = 0; c < C; c++) rho[c] = rho[c] + n[c][i] * R(GD[i]);
            const EdgeSink<R> es = edge_sink(edge, P, x, y, z);
#pragma unroll
            for (int c = 0; c < C; c++)
#pragma unroll
                for (int i = 0; i < 19; i++) es(c, i, n[c][i]);
            da[0] = da[1] = da[2] = R(0);
            return;
        }
    }
    bool self_rock = false;
    if constexpr (Rock) self_rock = rock[sidx(x, y, z, P.nx, P.ny)] != 0;
#pragma unroll
    for (int ii = 0; ii < 19; ii++) {
        const int i = dir_at<Dev && C == 3>(ii);
#pragma unroll
        for (int c = 0; c < C; c++) {
            R v;
            if constexpr (Rock)
                v = pull_rock<R, C>(f, rock, P, c, i, x, y, z, self_rock);
            else if constexpr (Flat)
                v = pull_flat<R, C>(f, P, c, i, x, y, z);
            else
                v = pull<R, C, LE>(f, face_pc, P, c, i, x, y, z);
            rho[c] = rho[c] + v * R(GD[i]);
            if constexpr (C == 3)
                if (c == 2) acc_da<R, Dev>(P, db, i, v, x, y, z, da0, da1, da2);
        }
    }
    da[0] = da0;
    da[1] = da1;
    da[2] = da2;
}

// PK (ternary, the batch's three-pass plan): the packed aux layout, one
// 4-vector (psi_r, psi_b, psi_s, rho_s) per site in place of the six planes,
// so that the force stencil loads one vector per neighbour
template <typename R, bool Dev, int C, bool LE, bool Rock = false, bool ZH = false, bool PK = false>
__device__ __forceinline__ void moments_site(const Phys<R> &P, const R *__restrict__ f,
                                             const R *__restrict__ face_pc,
                                             const R *__restrict__ db, R *__restrict__ aux,
                                             uint32_t s, const uint8_t *__restrict__ rock = nullptr,
                                             R *__restrict__ edge_adv = nullptr)
{
    int x, y, z;
    site_xyz(s, P.nx, P.ny, x, y, z);
    R rho[C], da[3];
    moments_core<R, Dev, C, LE, Rock, ZH>(P, f, face_pc, db, x, y, z, rho, da, rock, edge_adv);
    if constexpr (PK && C == 3) {
        ((V4<R> *)aux)[s] = V4<R>{psi_of<Dev>(rho[0], P.emr0[0]), psi_of<Dev>(rho[1], P.emr0[1]),
                                  psi_of<Dev>(rho[2], P.emr0[2]), rho[2]};
    } else {
#pragma unroll
        for (int c = 0; c < C; c++) {
            st_plane(aux + (size_t)c * P.ns, s, psi_of<Dev>(rho[c], P.emr0[c]));
            st_plane(aux + (size_t)(C + c) * P.ns, s, rho[c]);
        }
    }
    if constexpr (C == 3) {
        V4<R> *vec = (V4<R> *)(aux + aux3_vec_off(P.ns));
        vec[s] = V4<R>{rho[1] - rho[0], da[0], da[1], da[2]};
    }
}

template <typename R, bool Dev, int C, bool LE, bool Rock = false, bool ZH = false>
__global__ __launch_bounds__(256) void k_moments(Phys<R> P, const R *__restrict__ f,
                                                 const R *__restrict__ face_pc,
                                                 const R *__restrict__ db, R *__restrict__ aux,
                                                 uint32_t xg, const uint8_t *__restrict__ rock,
                                                 R *__restrict__ edge_adv = nullptr)
{
    uint32_t s = xcd_tile(blockIdx.x, gridDim.x, xg) * blockDim.x + threadIdx.x;
    if (s >= P.ns) return;
    moments_site<R, Dev, C, LE, Rock, ZH>(P, f, face_pc, db, aux, s, rock, edge_adv);
}

// Fills the edge columns of the pre-advection halo planes from the
// populations f (when no collide of the two-pass plan has written them).
template <typename R>
__global__ __launch_bounds__(256) void k_le_gather_edges(Phys<R> P, const R *__restrict__ f, R *__restrict__ edge_pc)
{
    const uint32_t plane = (uint32_t)P.ny * (uint32_t)P.nz;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;  // (col*plane + q)*38 + c*19 + i
    if (t >= 2u * 38u * plane) return;
    const uint32_t ci = t % 38u, q = (t / 38u) % plane, col = t / (38u * plane);
    const int y = (int)(q % (uint32_t)P.ny), z = (int)(q / (uint32_t)P.ny);
    const int x = col == 0 ? P.nx - 1 : 0;
    edge_pc[t] = f[ci * P.ns + sidx(x, y, z, P.nx, P.ny)];
}

// --------------------------------------------------------------------------
// c_i . (a0, a1, a2) over the nonzero components of c_i, in x, y, z order.  A
// zero component's product is an exact zero that leaves the sum unchanged
// (finite operands), so the terms are skipped instead of multiplied out; the
// -0.0 start is the additive identity the compiler folds away.
template <typename R>
__device__ __forceinline__ R dot_c(int i, R a0, R a1, R a2)
{
    R t = R(-0.0);
    if (CX[i] != 0) t = t + R(CX[i]) * a0;
    if (CY[i] != 0) t = t + R(CY[i]) * a1;
    if (CZ[i] != 0) t = t + R(CZ[i]) * a2;
    return t;
}

// --------------------------------------------------------------------------
// k_dipole: dipole relaxation, code/lb3d_collision.F90:296-403 (ternary).
// The dipole update of one site from its own advected dipole hs = da(x) and
// surfactant density, and per neighbour i (x + c_i) get_drb(i) = rho_b - rho_r
// and get_da(i, a).  Writes the new db(x) to out.
template <typename R, bool Dev, typename GetDrb, typename GetDa>
__device__ __forceinline__ void dipole_core(const Phys<R> &P, R hs0, R hs1, R hs2, R rho_s, GetDrb get_drb,
                                            GetDa get_da, R (&out)[3])
{
    if constexpr (Dev) rho_s = P.rho0[2] + rho_s;
    R sN = rho_s > R(CLAMP_SURF) ? rho_s : R(CLAMP_SURF);
    // FP64 divides as the reference does; the FP32 form multiplies by one
    // reciprocal per divisor (8 divisions -> 3 per site)
    R sd0, sd1, sd2;
    if constexpr (Dev) {
        const R rs = R(1.0) / sN;
        sd0 = hs0 * rs;
        sd1 = hs1 * rs;
        sd2 = hs2 * rs;
    } else {
        sd0 = hs0 / sN;
        sd1 = hs1 / sN;
        sd2 = hs2 / sN;
    }
    R hc0 = R(0), hc1 = R(0), hc2 = R(0);
#pragma unroll
    for (int i = 0; i < 18; i++) {
        // -sign(1,g_bs)*(rho_b - rho_r)(x+c_i)*g_i (raw densities, :361-369)
        R gs = R(GD[i]);
        R hc_m = -((P.sign_g_bs * get_drb(i)) * gs);
        if (CX[i] != 0) hc0 = hc0 + hc_m * R(CX[i]);
        if (CY[i] != 0) hc1 = hc1 + hc_m * R(CY[i]);
        if (CZ[i] != 0) hc2 = hc2 + hc_m * R(CZ[i]);
        R p0 = get_da(i, 0), p1 = get_da(i, 1), p2 = get_da(i, 2);
        R hs_dot = R(2.0) * dot_c<R>(i, p0, p1, p2);
        // (p - c_i hs_dot) g_i; a zero component of c_i leaves p
        R q0 = p0, q1 = p1, q2 = p2;
        if (CX[i] != 0) q0 = p0 - R(CX[i]) * hs_dot;
        if (CY[i] != 0) q1 = p1 - R(CY[i]) * hs_dot;
        if (CZ[i] != 0) q2 = p2 - R(CZ[i]) * hs_dot;
        hs0 = hs0 + q0 * gs;
        hs1 = hs1 + q1 * gs;
        hs2 = hs2 + q2 * gs;
    }
    R h0 = hc0 + hs0, h1 = hc1 + hs1, h2 = hc2 + hs2;
    R h_m = P.beta * sqrt(h0 * h0 + h1 * h1 + h2 * h2);
    R e0 = P.third_beta * h0, e1 = P.third_beta * h1, e2 = P.third_beta * h2;
    if (h_m > R(10.0e-3)) {
        R exh = R(1.0) / tanh(h_m);
        if constexpr (Dev) {
            const R rh = R(1.0) / h_m;
            const R sd_mag = exh - rh;
            e0 = ((P.beta * h0) * rh) * sd_mag;
            e1 = ((P.beta * h1) * rh) * sd_mag;
            e2 = ((P.beta * h2) * rh) * sd_mag;
        } else {
            R sd_mag = exh - R(1.0) / h_m;
            e0 = ((P.beta * h0) / h_m) * sd_mag;
            e1 = ((P.beta * h1) / h_m) * sd_mag;
            e2 = ((P.beta * h2) / h_m) * sd_mag;
        }
    }
    e0 = e0 * P.d_0;
    e1 = e1 * P.d_0;
    e2 = e2 * P.d_0;
    out[0] = P.omega_d * e0 + P.one_m_omega_d * sd0;
    out[1] = P.omega_d * e1 + P.one_m_omega_d * sd1;
    out[2] = P.omega_d * e2 + P.one_m_omega_d * sd2;
}

template <typename R, bool Dev, bool PK = false>
__device__ __forceinline__ void dipole_site(const Phys<R> &P, const R *__restrict__ aux,
                                            R *__restrict__ db, uint32_t s)
{
    int x, y, z;
    site_xyz(s, P.nx, P.ny, x, y, z);
    const V4<R> *vec = (const V4<R> *)(aux + aux3_vec_off(P.ns));
    // the 18 neighbours' (rho_b - rho_r, da): one vector load each, at a
    // 32-bit byte offset from the field's base (16 ns bytes < 4 GiB for any
    // ternary lattice that 32-bit population indexing admits)
    V4<R> nb[18];
#pragma unroll
    for (int i = 0; i < 18; i++)
        nb[i] = ld_plane(vec, sidx(wrap1(x + CX[i], P.nx), wrap1(y + CY[i], P.ny), wrap1(z + CZ[i], P.nz), P.nx, P.ny));
    const V4<R> own = ld_plane(vec, s);
    R out[3];
    dipole_core<R, Dev>(
        P, own.b, own.c, own.d, PK ? ld_plane((const V4<R> *)aux, s).d : ld_plane(plane_of(aux, 5u, P.ns), s),
        [&](int i) __attribute__((always_inline)) { return nb[i].a; },
        [&](int i, int a) __attribute__((always_inline)) {
            return a == 0 ? nb[i].b : (a == 1 ? nb[i].c : nb[i].d);
        },
        out);
    st_plane(db, s, out[0]);
    st_plane(db + (size_t)P.ns, s, out[1]);
    st_plane(db + (size_t)2 * P.ns, s, out[2]);
}

template <typename R, bool Dev>
__global__ __launch_bounds__(256) void k_dipole(Phys<R> P, const R *__restrict__ aux,
                                                R *__restrict__ db, uint32_t xg)
{
    uint32_t s = xcd_tile(blockIdx.x, gridDim.x, xg) * blockDim.x + threadIdx.x;
    if (s >= P.ns) return;
    dipole_site<R, Dev>(P, aux, db, s);
}

// Ternary Shan-Chen forces (code/lb3d_force_shanchen.F90:215-330) of one site
// from its dipole dbx and psi ps_own (stored form), and per neighbour i
// get_psi(i, c) (stored form) and get_db(i, a).
template <typename R, bool Dev, typename GetPsi, typename GetDb>
__device__ __forceinline__ void sc_forces3(const Phys<R> &P, const R (&dbx)[3], const R (&ps_own)[3],
                                           GetPsi get_psi, GetDb get_db, R F[3][3])
{
    R frb[3] = {R(0), R(0), R(0)}, fbr[3] = {R(0), R(0), R(0)};
    R fcs[3] = {R(0), R(0), R(0)}, fsc[3] = {R(0), R(0), R(0)}, fss[3] = {R(0), R(0), R(0)};
    const R dbx0 = dbx[0], dbx1 = dbx[1], dbx2 = dbx[2];
    const R dpsi_br = Dev ? P.psi0[1] - P.psi0[0] : R(0);
#pragma unroll
    for (int i = 0; i < 18; i++) {
        const R c0 = R(CX[i]), c1 = R(CY[i]), c2 = R(CZ[i]);
        const R gs = R(GD[i]);
        R pr = get_psi(i, 0);
        R pb = get_psi(i, 1);
        R ps = get_psi(i, 2);
        if constexpr (Dev) ps = P.psi0[2] + ps;
        R green_r = pb * gs, green_b = pr * gs;
        if (CX[i] != 0) { frb[0] = frb[0] + green_r * c0; fbr[0] = fbr[0] + green_b * c0; }
        if (CY[i] != 0) { frb[1] = frb[1] + green_r * c1; fbr[1] = fbr[1] + green_b * c1; }
        if (CZ[i] != 0) { frb[2] = frb[2] + green_r * c2; fbr[2] = fbr[2] + green_b * c2; }
        // a zero component of c_i makes its terms exact zeros, which are
        // skipped (dot_c and the conditionals); the rest keep the order
        R d0 = get_db(i, 0) * ps;
        R d1 = get_db(i, 1) * ps;
        R d2 = get_db(i, 2) * ps;
        R d_dotc = dot_c<R>(i, d0, d1, d2);
        fcs[0] = fcs[0] + (CX[i] != 0 ? d0 - (R(2.0) * d_dotc) * c0 : d0) * gs;
        fcs[1] = fcs[1] + (CY[i] != 0 ? d1 - (R(2.0) * d_dotc) * c1 : d1) * gs;
        fcs[2] = fcs[2] + (CZ[i] != 0 ? d2 - (R(2.0) * d_dotc) * c2 : d2) * gs;
        R di_dotc = dot_c<R>(i, dbx0, dbx1, dbx2);
        R greenc_s = ((pb - pr) + dpsi_br) * gs;
        fsc[0] = fsc[0] + greenc_s * (CX[i] != 0 ? dbx0 - (R(2.0) * di_dotc) * c0 : dbx0);
        fsc[1] = fsc[1] + greenc_s * (CY[i] != 0 ? dbx1 - (R(2.0) * di_dotc) * c1 : dbx1);
        fsc[2] = fsc[2] + greenc_s * (CZ[i] != 0 ? dbx2 - (R(2.0) * di_dotc) * c2 : dbx2);
        R di_dot_d = d0 * dbx0 + d1 * dbx1 + d2 * dbx2;
        R green_ss = R(8.0) * gs;
        R a = di_dot_d - (R(2.0) * di_dotc) * d_dotc;
        R q0 = d0 * di_dotc, q1 = d1 * di_dotc, q2 = d2 * di_dotc;
        if (CX[i] != 0) q0 = a * c0 + d0 * di_dotc;
        if (CY[i] != 0) q1 = a * c1 + d1 * di_dotc;
        if (CZ[i] != 0) q2 = a * c2 + d2 * di_dotc;
        fss[0] = fss[0] + green_ss * (q0 + d_dotc * dbx0);
        fss[1] = fss[1] + green_ss * (q1 + d_dotc * dbx1);
        fss[2] = fss[2] + green_ss * (q2 + d_dotc * dbx2);
    }
    R psr = ps_own[0], psb = ps_own[1], pss = ps_own[2];
    if constexpr (Dev) {
        psr = P.psi0[0] + psr;
        psb = P.psi0[1] + psb;
        pss = P.psi0[2] + pss;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        R f_br = -(P.g_br * fbr[k]);
        R f_rb = -(P.g_br * frb[k]);
        R f_sc = P.two_g_bs * fsc[k];
        R f_cs = -(P.two_g_bs * fcs[k]);
        R f_ss = -(P.g_ss * fss[k]);
        F[1][k] = (f_br + f_cs) * psb;
        F[0][k] = (f_rb - f_cs) * psr;
        F[2][k] = (f_sc + f_ss) * pss;
    }
}

// --------------------------------------------------------------------------
// Shan-Chen forces at site (x,y,z) from the aux field.
// binary: code/lb3d_force_shanchen.F90:127-210; ternary: :215-330.
// Binary Shan-Chen forces next to walls (code/lb3d_force_shanchen.F90:127-210):
// a rock neighbour enters the wall sums with psi(rho + rock colour) and the
// wall couplings g_wr, g_wb weighted by tau_w/tau; fluid neighbours as usual.
// psi values are used whole (FP32 form: psi0 + deviation), since the wall
// and fluid sums no longer run over all 18 directions.
template <typename R, bool Dev>
__device__ __forceinline__ void sc_forces_wall(const Phys<R> &P, const R *__restrict__ aux,
                                               const uint8_t *__restrict__ rock, int x, int y, int z,
                                               uint32_t s, R F[2][3])
{
    R frb[3] = {R(0), R(0), R(0)}, fbr[3] = {R(0), R(0), R(0)};
    R fwr[3] = {R(0), R(0), R(0)}, fwb[3] = {R(0), R(0), R(0)};
#pragma unroll
    for (int i = 0; i < 18; i++) {
        int xa = wrap1(x + CX[i], P.nx), ya = wrap1(y + CY[i], P.ny), za = wrap1(z + CZ[i], P.nz);
        uint32_t q = sidx(xa, ya, za, P.nx, P.ny);
        R pr, pb;
        const bool wall = rock[q] != 0;
        if (wall) {
            pr = psi_of<Dev>(ld_plane(plane_of(aux, 2u, P.ns), q) + P.rock_col[0], P.emr0[0]);
            pb = psi_of<Dev>(ld_plane(plane_of(aux, 3u, P.ns), q) + P.rock_col[1], P.emr0[1]);
        } else {
            pr = ld_plane(aux, q);
            pb = ld_plane(plane_of(aux, 1u, P.ns), q);
        }
        if constexpr (Dev) {
            pr = P.psi0[0] + pr;
            pb = P.psi0[1] + pb;
        }
        R green_r = pb * R(GD[i]), green_b = pr * R(GD[i]);
        R *sr = wall ? fwb : frb, *sb = wall ? fwr : fbr;
        if (CX[i] != 0) { sr[0] = sr[0] + green_r * R(CX[i]); sb[0] = sb[0] + green_b * R(CX[i]); }
        if (CY[i] != 0) { sr[1] = sr[1] + green_r * R(CY[i]); sb[1] = sb[1] + green_b * R(CY[i]); }
        if (CZ[i] != 0) { sr[2] = sr[2] + green_r * R(CZ[i]); sb[2] = sb[2] + green_b * R(CZ[i]); }
    }
    R psr = ld_plane(aux, s), psb = ld_plane(plane_of(aux, 1u, P.ns), s);
    if constexpr (Dev) {
        psr = P.psi0[0] + psr;
        psb = P.psi0[1] + psb;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        R wb = -(P.g_wb * fwb[k]), wr = -(P.g_wr * fwr[k]);
        F[1][k] = ((-(P.g_br * fbr[k])) + (wb / P.tau[1]) * P.tau_wb) * psb;
        F[0][k] = ((-(P.g_br * frb[k])) + (wr / P.tau[0]) * P.tau_wr) * psr;
    }
}

template <typename R, bool Dev, int C, bool LE, bool PK = false>
__device__ __forceinline__ void sc_forces(const Phys<R> &P, const R *__restrict__ aux,
                                          const R *__restrict__ face_aux,
                                          const R *__restrict__ db, int x, int y, int z,
                                          uint32_t s, R F[C][3])
{
    R frb[3] = {R(0), R(0), R(0)}, fbr[3] = {R(0), R(0), R(0)};
    if constexpr (C == 2) {
#pragma unroll
        for (int i = 0; i < 18; i++) {
            int xa = x + CX[i], ya = y + CY[i], za = z + CZ[i];
            R pr = aux_at<R, LE>(aux, face_aux, P, 0, xa, ya, za);
            R pb = aux_at<R, LE>(aux, face_aux, P, 1, xa, ya, za);
            R green_r = pb * R(GD[i]);
            R green_b = pr * R(GD[i]);
            if (CX[i] != 0) { frb[0] = frb[0] + green_r * R(CX[i]); fbr[0] = fbr[0] + green_b * R(CX[i]); }
            if (CY[i] != 0) { frb[1] = frb[1] + green_r * R(CY[i]); fbr[1] = fbr[1] + green_b * R(CY[i]); }
            if (CZ[i] != 0) { frb[2] = frb[2] + green_r * R(CZ[i]); fbr[2] = fbr[2] + green_b * R(CZ[i]); }
        }
        R psr = ld_plane(aux, s), psb = ld_plane(plane_of(aux, 1u, P.ns), s);
        if constexpr (Dev) {
            psr = P.psi0[0] + psr;
            psb = P.psi0[1] + psb;
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            F[1][k] = (-(P.g_br * fbr[k])) * psb;
            F[0][k] = (-(P.g_br * frb[k])) * psr;
        }
    } else if constexpr (PK) {
        const R dbx[3] = {ld_plane(db, s), ld_plane(plane_of(db, 1u, P.ns), s), ld_plane(plane_of(db, 2u, P.ns), s)};
        const V4<R> *ps4 = (const V4<R> *)aux;
        const V4<R> po = ld_plane(ps4, s);
        const R ps_own[3] = {po.a, po.b, po.c};
        V4<R> pn[18];
#pragma unroll
        for (int i = 0; i < 18; i++)
            pn[i] = ld_plane(ps4, sidx(wrap1(x + CX[i], P.nx), wrap1(y + CY[i], P.ny), wrap1(z + CZ[i], P.nz), P.nx, P.ny));
        sc_forces3<R, Dev>(
            P, dbx, ps_own,
            [&](int i, int c) __attribute__((always_inline)) {
                return c == 0 ? pn[i].a : (c == 1 ? pn[i].b : pn[i].c);
            },
            [&](int i, int a) __attribute__((always_inline)) {
                return aux_at<R, false>(db, nullptr, P, a, x + CX[i], y + CY[i], z + CZ[i]);
            },
            F);
    } else {
        const R dbx[3] = {ld_plane(db, s), ld_plane(plane_of(db, 1u, P.ns), s), ld_plane(plane_of(db, 2u, P.ns), s)};
        const R ps_own[3] = {ld_plane(aux, s), ld_plane(plane_of(aux, 1u, P.ns), s),
                             ld_plane(plane_of(aux, 2u, P.ns), s)};
        sc_forces3<R, Dev>(
            P, dbx, ps_own,
            [&](int i, int c) __attribute__((always_inline)) {
                return aux_at<R, false>(aux, nullptr, P, c, x + CX[i], y + CY[i], z + CZ[i]);
            },
            [&](int i, int a) __attribute__((always_inline)) {
                return aux_at<R, false>(db, nullptr, P, a, x + CX[i], y + CY[i], z + CZ[i]);
            },
            F);
    }
}

// The Shan-Chen force field of the state (the arr observable of the
// reference's dump_arrows, code/lb3d_io_dump_data.F90:1308-1460): the force
// of every site from the aux field of the last advection and the dipole
// field, with periodic x (the reference's dumps see periodic halos also under
// Lees-Edwards); rock sites feel none.  out[(c*3 + k)*ns + s].
template <typename R, bool Dev, int C, bool Rock>
__global__ __launch_bounds__(256) void k_forces(Phys<R> P, const R *__restrict__ aux, const R *__restrict__ db,
                                                const uint8_t *__restrict__ rock, R *__restrict__ out)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= P.ns) return;
    int x, y, z;
    site_xyz(s, P.nx, P.ny, x, y, z);
    R F[C][3];
    if constexpr (Rock) {
        if (rock[s]) {
#pragma unroll
            for (int c = 0; c < C; c++) F[c][0] = F[c][1] = F[c][2] = R(0);
        } else {
            sc_forces_wall<R, Dev>(P, aux, rock, x, y, z, s, F);
        }
    } else {
        sc_forces<R, Dev, C, false>(P, aux, nullptr, db, x, y, z, s, F);
    }
#pragma unroll
    for (int c = 0; c < C; c++)
#pragma unroll
        for (int k = 0; k < 3; k++) st_plane(out + (size_t)(c * 3 + k) * P.ns, s, F[c][k]);
}

// --------------------------------------------------------------------------
// Collision of one site from its streamed populations n and Shan-Chen forces
// F: moments, body force, BGK relaxation towards rho*boltz(u_tilde, du),
// code/lb3d_collision.F90:188-289 with lb3d_get_sc_u_eq (:560-587) and
// lb3d_get_sc_du (:523-556).  Writes the post-collision populations of site s.
// get(c, i) returns population i of component c; with Refetch the moments
// pass and the relaxation pass each call it anew (a compiler memory barrier
// between them keeps the loads from being shared), so that no population
// stays live in registers across the two passes.
// Density and momentum of component c from one population v of direction i,
// accumulated in direction order; every kernel that needs the moments of the
// streamed state sums them through this, so all plans round alike.
template <typename R>
__device__ __forceinline__ void acc_moments(int i, R v, R &rho, R (&mom)[3])
{
    rho = rho + v * R(GD[i]);
    R ng = v * R(GD[i]);
    if (CX[i] != 0) mom[0] = mom[0] + ng * R(CX[i]);
    if (CY[i] != 0) mom[1] = mom[1] + ng * R(CY[i]);
    if (CZ[i] != 0) mom[2] = mom[2] + ng * R(CZ[i]);
}

// a relaxation that only stores its populations
struct NoSink {
    template <typename R>
    __device__ __forceinline__ void operator()(int, int, R) const
    {
    }
};
template <typename R, bool Dev, int C, int BD, bool NT = false, typename Get, typename Sink = NoSink>
__device__ __forceinline__ void bgk_relax(const Phys<R> &P, Get get, const R (&rho)[C], const R (&mom)[C][3],
                                          R (&F)[C][3], int x, int y, int z, uint32_t s, R *__restrict__ fout,
                                          bool store = true, Sink sink = Sink());

template <typename R, bool Dev, int C, int BD, bool Refetch = false, bool NT = false, typename Get,
          typename Sink = NoSink>
__device__ __forceinline__ void bgk_site_g(const Phys<R> &P, Get get, R (&F)[C][3], int x, int y, int z,
                                           uint32_t s, R *__restrict__ fout, Sink sink = Sink())
{
    R rho[C], mom[C][3];
#pragma unroll
    for (int c = 0; c < C; c++) {
        rho[c] = R(0);
        mom[c][0] = mom[c][1] = mom[c][2] = R(0);
    }
#pragma unroll
    for (int ii = 0; ii < 19; ii++) {
        const int i = dir_at<Dev && C == 3>(ii);
#pragma unroll
        for (int c = 0; c < C; c++) acc_moments<R>(i, get(c, i), rho[c], mom[c]);
    }
    if constexpr (Refetch) asm volatile("" ::: "memory");
    bgk_relax<R, Dev, C, BD, NT>(P, get, rho, mom, F, x, y, z, s, fout, true, sink);
}

// The relaxation of bgk_site_g from the moments (rho, mom) of the streamed
// populations that get(c, i) returns.  Each post-collision population also
// goes to sink(c, i, value) once get(c, i) has been read; store = false skips
// the stores to fout.
template <typename R, bool Dev, int C, int BD, bool NT, typename Get, typename Sink>
__device__ __forceinline__ void bgk_relax(const Phys<R> &P, Get get, const R (&rho)[C], const R (&mom)[C][3],
                                          R (&F)[C][3], int x, int y, int z, uint32_t s, R *__restrict__ fout,
                                          bool store, Sink sink)
{
    R rhof[C];  // full densities
#pragma unroll
    for (int c = 0; c < C; c++) rhof[c] = Dev ? P.rho0[c] + rho[c] : rho[c];
    bool inbox = false;
    if (P.body) {
        int gx = x + 1, gy = y + 1, gz = z + 1;
        inbox = gz >= P.bmin[2] && gz <= P.bmax[2] && gx >= P.bmin[0] && gx <= P.bmax[0] && gy >= P.bmin[1] &&
                gy <= P.bmax[1];
        if (inbox) {
#pragma unroll
            for (int c = 0; c < C; c++) {
                R mass = P.amass[c] * rhof[c];
                F[c][2] = F[c][2] + mass * P.g_accn[2];
                F[c][1] = F[c][1] + mass * P.g_accn[1];
                F[c][0] = F[c][0] + mass * P.g_accn[0];
            }
        }
    }
    // u_tilde = sum amass*omega*P / max(sum amass*rho*omega, 1e-8)
    R pt0 = P.am_om[0] * mom[0][0], pt1 = P.am_om[0] * mom[0][1], pt2 = P.am_om[0] * mom[0][2];
    R rt = (P.amass[0] * rhof[0]) * P.omega[0];
#pragma unroll
    for (int c = 1; c < C; c++) {
        pt0 = pt0 + P.am_om[c] * mom[c][0];
        pt1 = pt1 + P.am_om[c] * mom[c][1];
        pt2 = pt2 + P.am_om[c] * mom[c][2];
        rt = rt + (P.amass[c] * rhof[c]) * P.omega[c];
    }
    rt = rt > R(CLAMP_RHO) ? rt : R(CLAMP_RHO);
    // FP64 divides as the reference does (lb3d_get_sc_u_eq, lb3d_get_sc_du);
    // the FP32 form multiplies by 1/rt and tau/(amass*rho): 1 + C divisions
    // per site instead of 3 + 4C
    R ux, uy, uz;
    if constexpr (Dev) {
        const R irt = R(1.0) / rt;
        ux = pt0 * irt;
        uy = pt1 * irt;
        uz = pt2 * irt;
    } else {
        ux = pt0 / rt;
        uy = pt1 / rt;
        uz = pt2 / rt;
    }
#pragma unroll
    for (int c = 0; c < C; c++) {
        R pon = R(1.0) / (rhof[c] > R(CLAMP_RHO) ? rhof[c] : R(CLAMP_RHO));
        R du0, du1, du2;
        if constexpr (Dev) {
            const R k = P.tau_am[c] * pon;
            du0 = k * F[c][0];
            du1 = k * F[c][1];
            du2 = k * F[c][2];
        } else {
            R tp = P.tau[c] * pon;
            du0 = (tp * F[c][0]) / P.amass[c];
            du1 = (tp * F[c][1]) / P.amass[c];
            du2 = (tp * F[c][2]) / P.amass[c];
        }
        R b[19];
        if constexpr (BD == BD_GENERAL) {
            // the other equilibrium forms; Guo's dd = (tau - 1/2) g_accn inside the
            // body-force box (lb3d_get_guo_force, code/lb3d_collision.F90:505-519),
            // 0 outside it (where the reference leaves dd unset)
            R dd0 = R(0), dd1 = R(0), dd2 = R(0);
            if (inbox) {
                R tm = P.tau[c] - R(0.5);
                dd0 = tm * P.g_accn[0];
                dd1 = tm * P.g_accn[1];
                dd2 = tm * P.g_accn[2];
            }
            boltz_general<R>(P.bdist, ux, uy, uz, du0, du1, du2, dd0, dd1, dd2, b);
            if constexpr (Dev) {
#pragma unroll
                for (int i = 0; i < 19; i++) b[i] = b[i] - (i < 18 ? R(1.0 / 36.0) : R(1.0 / 3.0));
            }
        }
        if constexpr (Dev) {
            if constexpr (BD != BD_GENERAL) boltz_dev(BD, ux, uy, uz, du0, du1, du2, b);
#pragma unroll
            for (int i = 0; i < 19; i++) {
                // deviation of rho*b_i from w_i*rho0: w_i*drho + rho*(b_i - w_i)
                R wi = i < 18 ? R(1.0 / 36.0) : R(1.0 / 3.0);
                R deq = wi * rho[c] + rhof[c] * b[i];
                R v = get(c, i);
                R nv = v - P.omega[c] * (v - deq);
                if (store) st_pop<NT>(fout + (size_t)((uint32_t)c * 19u + (uint32_t)i) * P.ns, s, nv);
                sink(c, i, nv);
            }
        } else {
            if constexpr (BD != BD_GENERAL) boltz_dist<R>(BD, ux, uy, uz, du0, du1, du2, b);
#pragma unroll
            for (int i = 0; i < 19; i++) {
                R neq = rho[c] * b[i];
                R v = get(c, i);
                R nv = v - P.omega[c] * (v - neq);
                if (store) st_pop<NT>(fout + (size_t)((uint32_t)c * 19u + (uint32_t)i) * P.ns, s, nv);
                sink(c, i, nv);
            }
        }
    }
}

template <typename R, bool Dev, int C, int BD, bool NT = false, typename Sink = NoSink>
__device__ __forceinline__ void bgk_site(const Phys<R> &P, const R (&n)[C][19], R (&F)[C][3],
                                         int x, int y, int z, uint32_t s, R *__restrict__ fout,
                                         Sink sink = Sink())
{
    bgk_site_g<R, Dev, C, BD, false, NT>(P, [&](int c, int i) __attribute__((always_inline)) { return n[c][i]; }, F, x, y,
                              z, s, fout, sink);
}

// --------------------------------------------------------------------------
// k_collide: pull, Shan-Chen forces from the aux field, BGK; the second pass
// of the two-pass step (k_moments then k_collide).
// NT: bit 0 loads, bit 1 stores the populations with the non-temporal hint
template <typename R, bool Dev, int C, bool LE, int BD, bool Rock = false, bool ZH = false, int NT = 0,
          bool PK = false>
__device__ __forceinline__ void collide_site(const Phys<R> &P, const R *__restrict__ f,
                                             const R *__restrict__ face_pc,
                                             const R *__restrict__ aux,
                                             const R *__restrict__ face_aux,
                                             const R *__restrict__ db, R *__restrict__ fout,
                                             uint32_t s, const uint8_t *__restrict__ rock = nullptr,
                                             R *__restrict__ edge_pc = nullptr)
{
    int x, y, z;
    site_xyz(s, P.nx, P.ny, x, y, z);
    R n[C][19];
    R F[C][3];
    if constexpr (Rock) {
        static_assert(C == 2 && !LE, "walls: periodic binary lattices");
        const bool self_rock = rock[s] != 0;
#pragma unroll
        for (int i = 0; i < 19; i++)
#pragma unroll
            for (int c = 0; c < C; c++) n[c][i] = pull_rock<R, C>(f, rock, P, c, i, x, y, z, self_rock);
        // only fluid sites feel the Shan-Chen force (code/lb3d_collision.F90:160-172)
        if (self_rock) {
#pragma unroll
            for (int c = 0; c < C; c++) F[c][0] = F[c][1] = F[c][2] = R(0);
        } else {
            sc_forces_wall<R, Dev>(P, aux, rock, x, y, z, s, F);
        }
    } else {
        // the force stencil first, then the pulls (a compiler barrier keeps
        // them after it): the populations are not live while the stencil's
        // loads are in flight, which leaves the registers to batch those loads
        sc_forces<R, Dev, C, LE, PK>(P, aux, face_aux, db, x, y, z, s, F);
        asm volatile("" ::: "memory");
#pragma unroll
        for (int i = 0; i < 19; i++)
#pragma unroll
            for (int c = 0; c < C; c++) n[c][i] = pull<R, C, LE, (NT & 1) != 0>(f, face_pc, P, c, i, x, y, z);
    }
    if constexpr (ZH) {
        if (z == 0 || z == P.nz - 1) {
            // Zou-He planes: the rebuilt populations relax towards the
            // equilibrium of the streamed state's moments, which the reference
            // keeps from the advection (code/lb3d_collision.F90:249-289, 560-587)
            R rho[C], mom[C][3];
#pragma unroll
            for (int c = 0; c < C; c++) {
                rho[c] = R(0);
                mom[c][0] = mom[c][1] = mom[c][2] = R(0);
            }
#pragma unroll
            for (int i = 0; i < 19; i++)
#pragma unroll
                for (int c = 0; c < C; c++) acc_moments<R>(i, n[c][i], rho[c], mom[c]);
#pragma unroll
            for (int c = 0; c < C; c++) zou_he_planes<R, Dev>(P, n[c], c, z);
            bgk_relax<R, Dev, C, BD, (NT & 2) != 0>(
                P, [&](int c, int i) __attribute__((always_inline)) { return n[c][i]; }, rho, mom, F, x, y, z, s,
                fout);
            return;
        }
    }
    if constexpr (LE && C == 2)  // the edge columns' post-collision populations
        bgk_site<R, Dev, C, BD, (NT & 2) != 0>(P, n, F, x, y, z, s, fout, edge_sink(edge_pc, P, x, y, z));
    else
        bgk_site<R, Dev, C, BD, (NT & 2) != 0>(P, n, F, x, y, z, s, fout);
}

// NT: the populations' loads and stores with the non-temporal hint, for
// lattices whose populations do not fit the memory-side cache (see ld_pop)
template <typename R, bool Dev, int C, bool LE, int BD, bool Rock = false, bool ZH = false, bool NT = false>
__global__ __launch_bounds__(256) void k_collide(Phys<R> P, const R *__restrict__ f,
                                                 const R *__restrict__ face_pc,
                                                 const R *__restrict__ aux,
                                                 const R *__restrict__ face_aux,
                                                 const R *__restrict__ db, R *__restrict__ fout,
                                                 uint32_t xg, const uint8_t *__restrict__ rock,
                                                 R *__restrict__ edge_pc = nullptr,
                                                 LEState *__restrict__ le_commit = nullptr, int le_adv = 0)
{
    // Lees-Edwards: the step's state, which its halo-plane kernels computed
    // from the stored one, replaces it (no kernel of the step reads it after
    // them)
    if constexpr (LE)
        if (le_commit && blockIdx.x == 0 && threadIdx.x == 0) *le_commit = le_advance(*le_commit, P.nz, le_adv);
    uint32_t s = xcd_tile(blockIdx.x, gridDim.x, xg) * blockDim.x + threadIdx.x;
    if (s >= P.ns) return;
    // measured on MI355X (FP32 and FP64): ternary loads and stores (128^3:
    // -6 to -8% per step), binary stores only (256^3 FP64: -2%)
    constexpr int nt = !NT ? 0 : (C == 3 ? 3 : 2);
    collide_site<R, Dev, C, LE, BD, Rock, ZH, nt>(P, f, face_pc, aux, face_aux, db, fout, s, rock, edge_pc);
}

// --------------------------------------------------------------------------
// k_fused2: the whole binary timestep in one pass.  A workgroup owns a TX x TY
// column tile and marches along z through [z0, z0+lz).  Each thread pulls the
// streamed populations of its site one plane ahead (plane k+1) and keeps them
// in registers (or LDS) while it collides plane k; the densities psi of plane k+1 for
// the tile and its one-site halo go into an LDS ring of three planes, the
// halo sites by a sum-only pull.  Nothing but the populations is read from or
// written to memory (Lees-Edwards halo columns read face_pc / face_aux).
// Each population is read once by its owner; halo pulls re-read neighbours'
// data, which adjacent tiles on the same XCD fetch at the same time.
template <typename R, bool Dev, bool LE>
__device__ __forceinline__ void psi_at(const Phys<R> &P, const R *__restrict__ f,
                                       const R *__restrict__ face_pc,
                                       const R *__restrict__ face_aux, int x, int y, int z,
                                       R &p0, R &p1)
{
    if (LE && (x < 0 || x >= P.nx)) {
        uint32_t plane = (uint32_t)P.ny * (uint32_t)P.nz;
        int fc = x < 0 ? 0 : 1;
        uint32_t q = (uint32_t)wrap1(z, P.nz) * (uint32_t)P.ny + (uint32_t)wrap1(y, P.ny);
        p0 = ld_plane(plane_of(face_aux, (uint32_t)(fc * 4 + 0), plane), q);
        p1 = ld_plane(plane_of(face_aux, (uint32_t)(fc * 4 + 1), plane), q);
        return;
    }
    int xs = wrap1(x, P.nx), ys = wrap1(y, P.ny), zs = wrap1(z, P.nz);
    R r0 = R(0), r1 = R(0);
#pragma unroll
    for (int i = 0; i < 19; i++) {
        r0 = r0 + pull<R, 2, LE>(f, face_pc, P, 0, i, xs, ys, zs) * R(GD[i]);
        r1 = r1 + pull<R, 2, LE>(f, face_pc, P, 1, i, xs, ys, zs) * R(GD[i]);
    }
    p0 = psi_of<Dev>(r0, P.emr0[0]);
    p1 = psi_of<Dev>(r1, P.emr0[1]);
}

// SM selects where the streamed populations of plane k wait for the collide:
// 0 registers (pulled one plane ahead), 1 an LDS stage (pulled one plane
// ahead), 2 nowhere: the collide pulls them again, once for the moments and
// once for the relaxation (L2 hits one plane after the psi pass read them),
// which leaves the fewest live registers.
template <typename R, bool Dev, bool LE, int TX, int TY, int SM>
__global__ __launch_bounds__(TX *TY) void k_fused2(Phys<R> P, const R *__restrict__ f,
                                                    const R *__restrict__ face_pc,
                                                    const R *__restrict__ face_aux,
                                                    R *__restrict__ fout, int lz, uint32_t xg)
{
    constexpr int T = TX * TY, LX = TX + 2, LY = TY + 2, LP = LX * LY;
    constexpr int NH = 2 * LX + 2 * TY;  // halo positions per plane
    // psi of the tile + halo, plane p in slot p - z0 + 1 (mod 4): the collide of
    // plane k reads planes k-1..k+1 while a thread that is done with it may
    // already fill plane k+2, so one barrier per plane suffices
    __shared__ R ring[4][2][LP];
    // populations of plane k+1: in LDS, one column per thread (SM 1), or in registers (SM 0)
    __shared__ R stage[SM == 1 ? 38 : 1][SM == 1 ? T : 1];
    R nxt[2][19];
    const int tid = threadIdx.x;
    const int lx = tid % TX, ly = tid / TX;
    // tile (bx, by, bz) of a gx x gy x gz grid, x fastest, placed by xcd_tile
    const uint32_t gx = ((uint32_t)P.nx + TX - 1) / TX, gy = ((uint32_t)P.ny + TY - 1) / TY;
    const uint32_t gz = ((uint32_t)P.nz + lz - 1) / (uint32_t)lz;
    const uint32_t tl = xcd_tile(blockIdx.x, gridDim.x, xg);
    if (tl >= gx * gy * gz) return;  // uniform over the block
    const int x0 = (int)(tl % gx) * TX, y0 = (int)((tl / gx) % gy) * TY, z0 = (int)(tl / (gx * gy)) * lz;
    const int x = x0 + lx, y = y0 + ly;
    const bool own = x < P.nx && y < P.ny;
    const int zend = min(z0 + lz, P.nz);
    if (z0 >= zend) return;  // uniform over the block
    // LDS positions this thread fills besides its own: up to NHS halo slots
    constexpr int NHS = (NH + T - 1) / T;  // halo positions per thread
    int hq[NHS], hgx[NHS], hgy[NHS];
    bool hv[NHS];
#pragma unroll
    for (int r = 0; r < NHS; r++) {
        int h = tid + r * T;
        int ax = 0, ay = 0;
        if (h < LX) { ax = h - 1; ay = -1; }
        else if (h < 2 * LX) { ax = h - LX - 1; ay = TY; }
        else if (h < 2 * LX + TY) { ax = -1; ay = h - 2 * LX; }
        else { ax = TX; ay = h - 2 * LX - TY; }
        hq[r] = (ay + 1) * LX + (ax + 1);
        hgx[r] = x0 + ax;
        hgy[r] = y0 + ay;
        // needed: at most one site beyond the lattice in x and y
        hv[r] = h < NH && x0 + ax <= P.nx && y0 + ay <= P.ny;
    }
    const bool mid_needed = !own && x <= P.nx && y <= P.ny;  // an interior slot acting as halo
    const int me = (ly + 1) * LX + (lx + 1);

    // psi of plane zp for the tile and halo into ring slot; the owner's
    // streamed populations of that plane go to its stage column
    auto fill = [&](int zp, int slot) __attribute__((always_inline)) {
        // opaque copies of the lane coordinates: keeps the compiler from
        // hoisting the 19 per-direction offsets out of the z loop (they
        // would stay live in VGPRs across the collide)
        int xo = x, yo = y;
        asm volatile("" : "+v"(xo), "+v"(yo));
        if (own) {
            int zs = wrap1(zp, P.nz);
            R r0 = R(0), r1 = R(0);
#pragma unroll
            for (int i = 0; i < 19; i++) {
                R v0 = pull<R, 2, LE>(f, face_pc, P, 0, i, xo, yo, zs);
                R v1 = pull<R, 2, LE>(f, face_pc, P, 1, i, xo, yo, zs);
                if constexpr (SM == 1) {
                    stage[i][tid] = v0;
                    stage[19 + i][tid] = v1;
                } else if constexpr (SM == 0) {
                    nxt[0][i] = v0;
                    nxt[1][i] = v1;
                }
                r0 = r0 + v0 * R(GD[i]);
                r1 = r1 + v1 * R(GD[i]);
            }
            ring[slot][0][me] = psi_of<Dev>(r0, P.emr0[0]);
            ring[slot][1][me] = psi_of<Dev>(r1, P.emr0[1]);
        } else if (mid_needed) {
            R p0, p1;
            psi_at<R, Dev, LE>(P, f, face_pc, face_aux, x, y, zp, p0, p1);
            ring[slot][0][me] = p0;
            ring[slot][1][me] = p1;
        }
#pragma unroll
        for (int r = 0; r < NHS; r++)
            if (hv[r]) {
                R p0, p1;
                int gx = hgx[r], gy = hgy[r];
                asm volatile("" : "+v"(gx), "+v"(gy));
                psi_at<R, Dev, LE>(P, f, face_pc, face_aux, gx, gy, zp, p0, p1);
                ring[slot][0][hq[r]] = p0;
                ring[slot][1][hq[r]] = p1;
            }
    };
    // prologue: planes z0-1 and z0
    fill(z0 - 1, 0);
    fill(z0, 1);
    for (int k = z0; k < zend; k++) {
        const int sm = (k - z0) & 3, s0 = (k - z0 + 1) & 3, sp = (k - z0 + 2) & 3;  // planes k-1, k, k+1
        // the populations of plane k from the stage (own column only)
        R cur[2][19];
        if constexpr (SM != 2) {
#pragma unroll
            for (int i = 0; i < 19; i++) {
                if constexpr (SM == 1) {
                    cur[0][i] = stage[i][tid];
                    cur[1][i] = stage[19 + i][tid];
                } else {
                    cur[0][i] = nxt[0][i];
                    cur[1][i] = nxt[1][i];
                }
            }
        }
        fill(k + 1, sp);
        __syncthreads();
        if (own) {
            R frb0 = R(0), frb1 = R(0), frb2 = R(0), fbr0 = R(0), fbr1 = R(0), fbr2 = R(0);
#pragma unroll
            for (int i = 0; i < 18; i++) {
                int sl = CZ[i] == 0 ? s0 : (CZ[i] > 0 ? sp : sm);
                int q = me + CY[i] * LX + CX[i];
                R pr = ring[sl][0][q];
                R pb = ring[sl][1][q];
                R green_r = pb * R(GD[i]);
                R green_b = pr * R(GD[i]);
                if (CX[i] != 0) { frb0 = frb0 + green_r * R(CX[i]); fbr0 = fbr0 + green_b * R(CX[i]); }
                if (CY[i] != 0) { frb1 = frb1 + green_r * R(CY[i]); fbr1 = fbr1 + green_b * R(CY[i]); }
                if (CZ[i] != 0) { frb2 = frb2 + green_r * R(CZ[i]); fbr2 = fbr2 + green_b * R(CZ[i]); }
            }
            R psr = ring[s0][0][me], psb = ring[s0][1][me];
            if constexpr (Dev) {
                psr = P.psi0[0] + psr;
                psb = P.psi0[1] + psb;
            }
            R F[2][3];
            F[1][0] = (-(P.g_br * fbr0)) * psb;
            F[1][1] = (-(P.g_br * fbr1)) * psb;
            F[1][2] = (-(P.g_br * fbr2)) * psb;
            F[0][0] = (-(P.g_br * frb0)) * psr;
            F[0][1] = (-(P.g_br * frb1)) * psr;
            F[0][2] = (-(P.g_br * frb2)) * psr;
            if constexpr (SM == 2)
                bgk_site_g<R, Dev, 2, 2, true>(
                    P,
                    [&](int c, int i) __attribute__((always_inline)) {
                        return pull<R, 2, LE>(f, face_pc, P, c, i, x, y, k);
                    },
                    F, x, y, k, sidx(x, y, k, P.nx, P.ny), fout);
            else
                bgk_site<R, Dev, 2, 2>(P, cur, F, x, y, k, sidx(x, y, k, P.nx, P.ny), fout);
        }
    }
}

// --------------------------------------------------------------------------
// The whole ternary timestep in one pass (periodic lattices): advection and
// moments (code/lb3d_advection.F90:54-151), dipole relaxation
// (code/lb3d_collision.F90:296-403), ternary Shan-Chen forces
// (code/lb3d_force_shanchen.F90:215-330) and BGK, marching along z through a
// TX x TY column tile.  The force at a site needs psi and the new dipole
// field one site around it, and the new dipole field needs rho_b - rho_r and
// the advected dipole one site further out.  Per plane k the work-group
//   M  computes the moments of plane k+2 on the tile and a ring of 2 sites:
//      each thread those of its own site, whose densities and momenta stay
//      in its registers until the collide of that plane, and the ring sites
//      spread over the threads;
//   D  relaxes the dipole field of plane k+1 on the tile and a ring of 1 site;
//   C  collides plane k from the registered moments, the forces and a
//      second pull of its populations.
// LDS keeps per site only what a later stage reads at that distance:
//   tile + 2 sites, 3 planes: rho_b - rho_r and the advected dipole;
//   tile + 1 site: psi of the 3 components (4 planes), rho_s (2 planes) and
//   the new dipole field (3 planes).
// The new dipole field of the tile's own sites goes to db_out; db_in (the
// previous step's field, read for the advected dipole) is left untouched, so
// neighbouring work-groups at other planes read consistent data.
template <int TX, int TY>
struct Fused3Dims {
    static constexpr int T = TX * TY;
    static constexpr int MX = TX + 4, MY = TY + 4, MP = MX * MY;  // tile + 2 sites
    static constexpr int DX = TX + 2, DY = TY + 2, DP = DX * DY;  // tile + 1 site
    static constexpr int NH = MP - T;                             // ring sites of M
    static constexpr int NHT = (NH + T - 1) / T;                  // ring sites per thread
    template <typename R>
    static constexpr size_t lds_bytes()
    {
        return (size_t)(3 * 4 * MP + (4 * 3 + 2 + 3 * 3) * DP) * sizeof(R);
    }
    // ring site h in [0, NH) of the moments region: rows 0-1, rows MY-2 and
    // MY-1, then columns 0, 1, MX-2, MX-1 of the rows between
    __device__ static __forceinline__ void ring_site(int h, int &mx, int &my)
    {
        if (h < 2 * MX) {
            my = h / MX;
            mx = h - my * MX;
        } else if (h < 4 * MX) {
            int g = h - 2 * MX;
            my = MY - 2 + g / MX;
            mx = g - (g / MX) * MX;
        } else {
            int g = h - 4 * MX;
            my = 2 + g / 4;
            int c = g & 3;
            mx = c < 2 ? c : MX - 4 + c;
        }
    }
};

template <typename R, bool Dev, int TX, int TY, int BD>
__global__ __launch_bounds__(TX *TY) void k_fused3(
    Phys<R> P, const R *__restrict__ f, const R *__restrict__ db_in, R *__restrict__ db_out,
    R *__restrict__ fout, int lz, uint32_t xg)
{
    using D = Fused3Dims<TX, TY>;
    __shared__ R aring[3][4][D::MP];  // rho_b - rho_r, da_xyz
    __shared__ R pring[4][3][D::DP];  // psi_r, psi_b, psi_s
    __shared__ R sring[2][D::DP];     // rho_s
    __shared__ R dring[3][3][D::DP];  // new dipole field
    const uint32_t gx = ((uint32_t)P.nx + TX - 1) / TX, gy = ((uint32_t)P.ny + TY - 1) / TY;
    const uint32_t gz = ((uint32_t)P.nz + lz - 1) / (uint32_t)lz;
    const uint32_t tl = xcd_tile(blockIdx.x, gridDim.x, xg);
    if (tl >= gx * gy * gz) return;  // uniform over the block
    const int x0 = (int)(tl % gx) * TX, y0 = (int)((tl / gx) % gy) * TY, z0 = (int)(tl / (gx * gy)) * lz;
    const int zend = min(z0 + lz, P.nz);
    if (z0 >= zend) return;  // uniform over the block

    const int tid = threadIdx.x;
    const int lx = tid % TX, ly = tid / TX;
    const int x = x0 + lx, y = y0 + ly;
    const bool own = x < P.nx && y < P.ny;
    // a tile slot past the lattice edge that the wrap still needs as ring
    const bool own_ring = !own && x <= P.nx + 1 && y <= P.ny + 1;
    const int qd = (ly + 1) * D::DX + (lx + 1);
    auto aslot = [&](int zp) __attribute__((always_inline)) { return (zp - z0 + 3) % 3; };
    auto pslot = [&](int zp) __attribute__((always_inline)) { return (zp - z0 + 4) & 3; };
    auto sslot = [&](int zp) __attribute__((always_inline)) { return (zp - z0 + 4) & 1; };
    auto dslot = [&](int zp) __attribute__((always_inline)) { return (zp - z0 + 3) % 3; };

    // the moments of one site into the rings at moments-region point (mx, my)
    auto put = [&](int zp, int mx, int my, const R (&rho)[3], const R (&da)[3]) __attribute__((always_inline)) {
        const int q = my * D::MX + mx, sa = aslot(zp);
        aring[sa][0][q] = rho[1] - rho[0];
        aring[sa][1][q] = da[0];
        aring[sa][2][q] = da[1];
        aring[sa][3][q] = da[2];
        if (mx >= 1 && mx <= D::MX - 2 && my >= 1 && my <= D::MY - 2) {
            const int p = (my - 1) * D::DX + (mx - 1), sp = pslot(zp);
            pring[sp][0][p] = psi_of<Dev>(rho[0], P.emr0[0]);
            pring[sp][1][p] = psi_of<Dev>(rho[1], P.emr0[1]);
            pring[sp][2][p] = psi_of<Dev>(rho[2], P.emr0[2]);
            sring[sslot(zp)][p] = rho[2];
        }
    };
    // M: moments of plane zp; the own site's densities and momenta to (hr, hm)
    auto fillM = [&](int zp, R (&hr)[3], R (&hm)[3][3]) __attribute__((always_inline)) {
        const int zw = wrap1(zp, P.nz);
        if (own) {
            // opaque copies of the lane coordinates: keeps the compiler from
            // hoisting the per-direction offsets out of the z loop
            int xo = x, yo = y;
            asm volatile("" : "+v"(xo), "+v"(yo));
            R da0 = R(0), da1 = R(0), da2 = R(0);
#pragma unroll
            for (int c = 0; c < 3; c++) {
                hr[c] = R(0);
                hm[c][0] = hm[c][1] = hm[c][2] = R(0);
            }
#pragma unroll
            for (int ii = 0; ii < 19; ii++) {
                const int i = dir_at<Dev>(ii);
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    R v = pull_flat<R, 3>(f, P, c, i, xo, yo, zw);
                    acc_moments<R>(i, v, hr[c], hm[c]);
                    if (c == 2) acc_da<R, Dev>(P, db_in, i, v, xo, yo, zw, da0, da1, da2);
                }
            }
            const R da[3] = {da0, da1, da2};
            put(zp, lx + 2, ly + 2, hr, da);
        } else if (own_ring) {
            int xo = wrap1(x, P.nx), yo = wrap1(y, P.ny);
            asm volatile("" : "+v"(xo), "+v"(yo));
            R rho[3], da[3];
            moments_core<R, Dev, 3, false, false, false, true>(P, f, nullptr, db_in, xo, yo, zw, rho, da);
            put(zp, lx + 2, ly + 2, rho, da);
        }
#pragma unroll
        for (int r = 0; r < D::NHT; r++) {
            const int h = tid + r * D::T;
            if (h >= D::NH) break;
            int mx, my;
            D::ring_site(h, mx, my);
            const int gxs = x0 - 2 + mx, gys = y0 - 2 + my;
            if (gxs > P.nx + 1 || gys > P.ny + 1) continue;  // beyond a partial tile's needs
            int xo = wrap1(gxs, P.nx), yo = wrap1(gys, P.ny);
            asm volatile("" : "+v"(xo), "+v"(yo));
            R rho[3], da[3];
            moments_core<R, Dev, 3, false, false, false, true>(P, f, nullptr, db_in, xo, yo, zw, rho, da);
            put(zp, mx, my, rho, da);
        }
    };
    // D: dipole field of plane zp on the tile + 1 site; own sites of planes
    // z0..zend-1 also go to db_out
    auto fillD = [&](int zp) __attribute__((always_inline)) {
        const int dsl = dslot(zp), ss = sslot(zp);
        const int sa0 = aslot(zp), sam = aslot(zp - 1), sap = aslot(zp + 1);
        for (int p = tid; p < D::DP; p += D::T) {
            const int dx = p % D::DX, dy = p / D::DX;
            const int gxs = x0 - 1 + dx, gys = y0 - 1 + dy;
            if (gxs > P.nx || gys > P.ny) continue;
            const int q = (dy + 1) * D::MX + (dx + 1);
            R out[3];
            dipole_core<R, Dev>(
                P, aring[sa0][1][q], aring[sa0][2][q], aring[sa0][3][q], sring[ss][p],
                [&](int i) __attribute__((always_inline)) {
                    int sl = CZ[i] == 0 ? sa0 : (CZ[i] > 0 ? sap : sam);
                    return aring[sl][0][q + CY[i] * D::MX + CX[i]];
                },
                [&](int i, int a) __attribute__((always_inline)) {
                    int sl = CZ[i] == 0 ? sa0 : (CZ[i] > 0 ? sap : sam);
                    return aring[sl][1 + a][q + CY[i] * D::MX + CX[i]];
                },
                out);
            dring[dsl][0][p] = out[0];
            dring[dsl][1][p] = out[1];
            dring[dsl][2][p] = out[2];
            if (dx >= 1 && dx <= TX && dy >= 1 && dy <= TY && gxs < P.nx && gys < P.ny && zp >= z0 && zp < zend) {
                uint32_t sg = sidx(gxs, gys, zp, P.nx, P.ny);
                st_plane(db_out, sg, out[0]);
                st_plane(db_out + (size_t)P.ns, sg, out[1]);
                st_plane(db_out + (size_t)2 * P.ns, sg, out[2]);
            }
        }
    };

    // densities and momenta of the own site at planes t-2, t-1, t
    R r0[3], m0[3][3], r1[3], m1[3][3], r2[3], m2[3][3];
    // iteration t: M of plane t, D of plane t-1, C of plane t-2; the first
    // iterations fill the rings (M from z0-2, D from z0-1)
    for (int t = z0 - 2; t < zend + 2; t++) {
        fillM(t, r2, m2);
        __syncthreads();
        if (t > z0 - 1) fillD(t - 1);
        __syncthreads();
        const int k = t - 2;
        if (own && k >= z0) {
            const int sp0 = pslot(k), spm = pslot(k - 1), spp = pslot(k + 1);
            const int sd0 = dslot(k), sdm = dslot(k - 1), sdp = dslot(k + 1);
            const R dbx[3] = {dring[sd0][0][qd], dring[sd0][1][qd], dring[sd0][2][qd]};
            const R ps_own[3] = {pring[sp0][0][qd], pring[sp0][1][qd], pring[sp0][2][qd]};
            R F[3][3];
            sc_forces3<R, Dev>(
                P, dbx, ps_own,
                [&](int i, int c) __attribute__((always_inline)) {
                    int sl = CZ[i] == 0 ? sp0 : (CZ[i] > 0 ? spp : spm);
                    return pring[sl][c][qd + CY[i] * D::DX + CX[i]];
                },
                [&](int i, int a) __attribute__((always_inline)) {
                    int sl = CZ[i] == 0 ? sd0 : (CZ[i] > 0 ? sdp : sdm);
                    return dring[sl][a][qd + CY[i] * D::DX + CX[i]];
                },
                F);
            int xo = x, yo = y;
            asm volatile("" : "+v"(xo), "+v"(yo));
            // the site's populations pulled again, all at once before the
            // relaxation (its stores would otherwise sit between the loads)
            R n[3][19];
#pragma unroll
            for (int i = 0; i < 19; i++)
#pragma unroll
                for (int c = 0; c < 3; c++) n[c][i] = pull_flat<R, 3>(f, P, c, i, xo, yo, k);
            bgk_relax<R, Dev, 3, BD>(
                P, [&](int c, int i) __attribute__((always_inline)) { return n[c][i]; }, r0, m0, F, x, y, k,
                sidx(x, y, k, P.nx, P.ny), fout);
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 3; c++) {
            r0[c] = r1[c];
            r1[c] = r2[c];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                m0[c][a] = m1[c][a];
                m1[c][a] = m2[c][a];
            }
        }
    }
}

// --------------------------------------------------------------------------
// Lees-Edwards x-halo planes, code/lb3d_bc_leesedwards.F90:659-1004 on one
// rank (all partners are self).  For (y,z) of the halo plane, with
// b = 0 (lelo >= 0) or -1 (lelo < 0) and f = lefrac:
//   low  (x=0):    (1-f) N(nx, y, z+lelo+b)   + f N(nx, y, z+lelo+1+b)
//   high (x=nx+1): f N(1, y, z-lelo-1-b)      + (1-f) N(1, y, z-lelo-b)
// (z wrapped), then n += rho_c*(feq(u -+ 2U) - feq(u)) with the negativity
// clamp; the high plane's clamp reads the finished low-plane site at the same
// (y,z) (:951-990).  Post = false: the sources are the post-collision
// populations (stage 6), written to face_pc.  Post = true: the sources are
// the post-advection populations (pulled through face_pc, stage 8) and only
// the halo moments are written to face_aux.
template <typename R, bool Dev, bool Post>
__device__ __forceinline__ R le_src(const Phys<R> &P, const R *__restrict__ f,
                                    const R *__restrict__ face_pc, int c, int i, int x, int y, int z)
{
    if constexpr (Post)
        return pull<R, 2, true>(f, face_pc, P, c, i, x, y, z);
    else
        return f[((uint32_t)c * 19u + (uint32_t)i) * P.ns + sidx(x, y, z, P.nx, P.ny)];
}

template <typename R, bool Dev>
__device__ __forceinline__ void le_correct(const Phys<R> &P, const LEState &L, R n[2][19],
                                           const R base[2][19], bool low)
{
    // moments of the interpolated halo site
    R rN[2], p[2][3];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        rN[c] = R(0);
        p[c][0] = p[c][1] = p[c][2] = R(0);
    }
#pragma unroll
    for (int i = 0; i < 19; i++)
#pragma unroll
        for (int c = 0; c < 2; c++) {
            rN[c] = rN[c] + n[c][i] * R(GD[i]);
            R ng = n[c][i] * R(GD[i]);
            if (CX[i] != 0) p[c][0] = p[c][0] + ng * R(CX[i]);
            if (CY[i] != 0) p[c][1] = p[c][1] + ng * R(CY[i]);
            if (CZ[i] != 0) p[c][2] = p[c][2] + ng * R(CZ[i]);
        }
    R rf[2];
#pragma unroll
    for (int c = 0; c < 2; c++) rf[c] = Dev ? P.rho0[c] + rN[c] : rN[c];
    R pt0 = P.am_om[0] * p[0][0], pt1 = P.am_om[0] * p[0][1], pt2 = P.am_om[0] * p[0][2];
    R rt = (P.amass[0] * rf[0]) * P.omega[0];
    pt0 = pt0 + P.am_om[1] * p[1][0];
    pt1 = pt1 + P.am_om[1] * p[1][1];
    pt2 = pt2 + P.am_om[1] * p[1][2];
    rt = rt + (P.amass[1] * rf[1]) * P.omega[1];
    rt = rt > R(CLAMP_RHO) ? rt : R(CLAMP_RHO);
    R ux = pt0 / rt, uy = pt1 / rt, uz = pt2 / rt;
    R su = R(L.shear_u);
    R uzs = low ? (uz - su) - su : (uz + su) + su;
    R b1[19], b2[19];
    if constexpr (Dev) {
        boltz_dev(P.bdist, ux, uy, uzs, R(0), R(0), R(0), b1);
        boltz_dev(P.bdist, ux, uy, uz, R(0), R(0), R(0), b2);
    } else {
        boltz_dist<R>(P.bdist, ux, uy, uzs, R(0), R(0), R(0), b1);
        boltz_dist<R>(P.bdist, ux, uy, uz, R(0), R(0), R(0), b2);
    }
    R df[2][19];
    bool neg = false;
#pragma unroll
    for (int i = 0; i < 19; i++) {
        R bd = b1[i] - b2[i];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            df[c][i] = rf[c] * bd;
            R nn = n[c][i] + df[c][i];
            R full = Dev ? (i < 18 ? P.wrho0_move[c] : P.wrho0_rest[c]) + nn : nn;
            if (full < R(0)) neg = true;
        }
    }
    if (neg) {
        bool allpos = true;
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int i = 0; i < 19; i++) {
                R full = Dev ? (i < 18 ? P.wrho0_move[c] : P.wrho0_rest[c]) + base[c][i] : base[c][i];
                if (!(full > R(0))) allpos = false;
            }
        R scale = R(0);
        if (allpos) {
            bool first = true;
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int i = 0; i < 19; i++) {
                    R full = Dev ? (i < 18 ? P.wrho0_move[c] : P.wrho0_rest[c]) + base[c][i] : base[c][i];
                    R v = -df[c][i] / full;
                    if (first || v > scale) scale = v;
                    first = false;
                }
        }
        bool use = fabs(scale) > R(1.0);
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int i = 0; i < 19; i++) n[c][i] = use ? base[c][i] + df[c][i] / scale : base[c][i];
    } else {
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int i = 0; i < 19; i++) n[c][i] = n[c][i] + df[c][i];
    }
}

__device__ __forceinline__ int wrapz1(int k, int n)
{  // k in 1-based terms, result 0-based index of wrap over 1..n
    int m = (k - 1) % n;
    if (m < 0) m += n;
    return m;
}

// One work-group of 2 * LE_SITES threads serves LE_SITES halo columns
// (y, z): its first wave builds the x = 0 plane's sites, the second the
// x = nx+1 plane's, both interpolating at once; the high side's clamp then
// reads the finished low site from LDS.
constexpr int LE_SITES = 64;
// Compact: the sources come from the edge columns (edge: the pre-advection
// ones, or for Post the streamed ones) instead of the lattice.
// adv: LE_STATE_AS_IS, or the step's state is le_advance(*Lp, nz, adv) (the
// collide of the step stores it; every block computes it alike)
template <typename R, bool Dev, bool Post, bool Compact = false>
__global__ __launch_bounds__(2 * LE_SITES) void k_le_faces(Phys<R> P, const LEState *__restrict__ Lp,
                                                           const R *__restrict__ f, R *__restrict__ face_pc,
                                                           R *__restrict__ face_aux,
                                                           const R *__restrict__ edge, int adv)
{
    __shared__ R lo_done[2][19][LE_SITES];  // the corrected low-plane sites
    const uint32_t plane = (uint32_t)P.ny * (uint32_t)P.nz;
    const int j = threadIdx.x % LE_SITES;
    const bool low = threadIdx.x < LE_SITES;  // wave-uniform
    const uint32_t t = blockIdx.x * LE_SITES + j;
    const bool live = t < plane;
    const LEState L = adv == LE_STATE_AS_IS ? *Lp : le_advance(*Lp, P.nz, adv);
    const int y = live ? (int)(t % (uint32_t)P.ny) : 0, z = live ? (int)(t / (uint32_t)P.ny) : 0;  // 0-based
    const int z1 = z + 1;                                                                           // 1-based
    const int lelo = L.lelo, bsh = lelo >= 0 ? 0 : -1;
    const R fr = R(L.lefrac), omf = R(1.0) - R(L.lefrac);
    R n[2][19];
    if (live) {
        // low (x=0):    (1-f) N(nx, y, z+lelo+b) + f N(nx, y, z+lelo+1+b)
        // high (x=nx+1): f N(1, y, z-lelo-1-b)   + (1-f) N(1, y, z-lelo-b)
        const int xs = low ? P.nx - 1 : 0;
        const int za = low ? wrapz1(z1 + lelo + bsh, P.nz) : wrapz1(z1 - lelo - 1 - bsh, P.nz);
        const int zb = low ? wrapz1(z1 + lelo + 1 + bsh, P.nz) : wrapz1(z1 - lelo - bsh, P.nz);
        const R wa = low ? omf : fr, wb = low ? fr : omf;
        const int col = low ? 0 : 1;
        const uint32_t qa = (uint32_t)za * (uint32_t)P.ny + (uint32_t)y, qb = (uint32_t)zb * (uint32_t)P.ny + (uint32_t)y;
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int i = 0; i < 19; i++) {
                R a, b;
                if constexpr (Compact) {
                    a = edge_site(edge, P, col, qa)[c * 19 + i];
                    b = edge_site(edge, P, col, qb)[c * 19 + i];
                } else {
                    a = le_src<R, Dev, Post>(P, f, face_pc, c, i, xs, y, za);
                    b = le_src<R, Dev, Post>(P, f, face_pc, c, i, xs, y, zb);
                }
                n[c][i] = wa * a + wb * b;
            }
        if (low) {
            R base[2][19];
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int i = 0; i < 19; i++) base[c][i] = n[c][i];
            le_correct<R, Dev>(P, L, n, base, true);
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int i = 0; i < 19; i++) lo_done[c][i][j] = n[c][i];
        }
    }
    __syncthreads();
    if (!live) return;
    if (!low) {
        R base[2][19];
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int i = 0; i < 19; i++) base[c][i] = lo_done[c][i][j];
        le_correct<R, Dev>(P, L, n, base, false);
    }
    const int fc = low ? 0 : 1;
    if constexpr (!Post) {
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int i = 0; i < 19; i++) face_pc[((uint32_t)(fc * 2 + c) * 19u + (uint32_t)i) * plane + t] = n[c][i];
    } else {
#pragma unroll
        for (int c = 0; c < 2; c++) {
            R rho = R(0);
#pragma unroll
            for (int i = 0; i < 19; i++) rho = rho + n[c][i] * R(GD[i]);
            face_aux[(uint32_t)(fc * 4 + c) * plane + t] = psi_of<Dev>(rho, P.emr0[c]);
            face_aux[(uint32_t)(fc * 4 + 2 + c) * plane + t] = rho;
        }
    }
}

// le_neighbours (code/lb3d_bc_leesedwards.F90:86-140, inv_fluid = 5) with the
// stage-6 update shear_u = shear_tmp*cos(shear_omega*nt) (code/lb3d_bc.F90:110),
// in FP64 on the device so that the step sequence can be replayed from a graph.
// --------------------------------------------------------------------------
// Batched ensemble: many independent periodic lattices of uneven sizes in the
// same launches.  Each lattice has its own parameters and a slice of the
// shared buffers; the work list holds (lattice, first site) tiles of 256
// sites.  Persistent grid: a fixed number of workgroups (a multiple of 8)
// strides over the list in xcd_tile order, which keeps each XCD on its own
// chunks of the list.
template <typename R>
struct BatchLat {
    Phys<R> P;
    uint64_t f_off, aux_off, db_off;  // element offsets of this lattice's slices
};

__device__ __forceinline__ uint2 batch_item(const uint2 *__restrict__ items, uint32_t it)
{
    uint2 w = items[it];
    w.x = __builtin_amdgcn_readfirstlane(w.x);
    w.y = __builtin_amdgcn_readfirstlane(w.y);
    return w;
}

// The batch's collide launches one kernel per equilibrium form present: the
// two compile-time forms bdist 0 and 2, and BD_GENERAL for the others.
template <int BD>
__device__ __forceinline__ bool batch_form_is(int bdist)
{
    if constexpr (BD == BD_GENERAL)
        return bdist != 0 && bdist != 2;
    else
        return bdist == BD;
}

template <typename R, bool Dev, int C, bool PK = false>
__global__ __launch_bounds__(256) void k_batch_moments(const BatchLat<R> *__restrict__ lat,
                                                       const uint2 *__restrict__ items,
                                                       uint32_t nitems, const R *__restrict__ f,
                                                       const R *__restrict__ db,
                                                       R *__restrict__ aux, uint32_t xg)
{
    const uint32_t n8 = xcd_round(nitems);
    for (uint32_t t = blockIdx.x; t < n8; t += gridDim.x) {
        uint32_t it = xcd_tile(t, n8, xg);
        if (it >= nitems) continue;
        uint2 w = batch_item(items, it);
        const BatchLat<R> &L = lat[w.x];
        uint32_t s = w.y + threadIdx.x;
        if (s < L.P.ns)
            moments_site<R, Dev, C, false, false, false, PK>(L.P, f + L.f_off, nullptr,
                                                             C == 3 ? db + L.db_off : nullptr, aux + L.aux_off, s);
    }
}

template <typename R, bool Dev, bool PK = false>
__global__ __launch_bounds__(256) void k_batch_dipole(const BatchLat<R> *__restrict__ lat,
                                                      const uint2 *__restrict__ items,
                                                      uint32_t nitems, const R *__restrict__ aux,
                                                      R *__restrict__ db, uint32_t xg)
{
    const uint32_t n8 = xcd_round(nitems);
    for (uint32_t t = blockIdx.x; t < n8; t += gridDim.x) {
        uint32_t it = xcd_tile(t, n8, xg);
        if (it >= nitems) continue;
        uint2 w = batch_item(items, it);
        const BatchLat<R> &L = lat[w.x];
        uint32_t s = w.y + threadIdx.x;
        if (s < L.P.ns) dipole_site<R, Dev, PK>(L.P, aux + L.aux_off, db + L.db_off, s);
    }
}

template <typename R, bool Dev, int C, int BD, bool NT, bool PK = false>
__global__ __launch_bounds__(256) void k_batch_collide(const BatchLat<R> *__restrict__ lat,
                                                       const uint2 *__restrict__ items,
                                                       uint32_t nitems, const R *__restrict__ f,
                                                       const R *__restrict__ aux,
                                                       const R *__restrict__ db,
                                                       R *__restrict__ fout, uint32_t xg)
{
    const uint32_t n8 = xcd_round(nitems);
    for (uint32_t t = blockIdx.x; t < n8; t += gridDim.x) {
        uint32_t it = xcd_tile(t, n8, xg);
        if (it >= nitems) continue;
        uint2 w = batch_item(items, it);
        const BatchLat<R> &L = lat[w.x];
        uint32_t s = w.y + threadIdx.x;
        // non-temporal stores of the populations (NT): -2% per step of the
        // 48-lattice ternary ensemble; non-temporal loads were slower
        if (s < L.P.ns && batch_form_is<BD>(L.P.bdist))
            collide_site<R, Dev, C, false, BD, false, false, NT ? 2 : 0, PK>(L.P, f + L.f_off, nullptr, aux + L.aux_off, nullptr,
                                                               C == 3 ? db + L.db_off : nullptr, fout + L.f_off, s);
    }
}

// --------------------------------------------------------------------------
// The collide-moments plan of the batched ternary ensemble.  A step is three
// launches: k_batch_dipole (the relaxed dipole field, from the moments aux_in
// of the step's streamed state), k_batch_cm (the collision, which also
// gathers the moments of the next step's streamed state from the
// post-collision populations it has just computed) and
// k_batch_edge_moments (the moments of the sites that k_batch_cm cannot
// gather).  The populations are read once per step (the three-pass plan
// reads them twice: once for the moments, once for the collision), and the
// moments of the streamed state are carried from one step to the next in a
// second aux field.
//
// Work item of k_batch_cm (uint4): x = lattice, y = x0 | y0 << 16,
// z = z0 | z1 << 16, w = tw | th << 9 | S << 18: a tile of tw columns and th
// rows of the lattice from (x0, y0), planes z0 .. z1-1, S >= tw + 2 the row
// stride of the work-group's region (S * th <= 256).
// Thread t holds region column rx = t % S, row ry = t / S, lattice site
// (x0 - 1 + rx (wrapped), y0 + ry).  Columns 1..tw are the tile's own;
// columns 0 and tw + 1 belong to the x neighbours and are collided again
// here (not stored), so that the post-collision populations that stream into
// the own columns are all at hand.  For each plane k = z0 .. z1-1:
//   1. collide the site of plane k (Shan-Chen forces from aux_in and the
//      relaxed dipole field db, pull, BGK; code/lb3d_collision.F90:122-289);
//      own columns store the populations to fout;
//   2. the post-collision populations of the 16 directions with an in-plane
//      component, and the site's dipole field, go to LDS;
//   3. each own site of rows 1..th-2 gathers what streams into it from plane
//      k (code/lb3d_advection.F90:54-151): directions with c_z = +1 into its
//      site of plane k+1, 0 into plane k, -1 into plane k-1, each group in
//      ascending order (ZORD), accumulating the density of each component and
//      the advected dipole da = sum_i g_i n_s,i db(x - c_i);
//   4. plane k-1 has then received all 19 directions; for planes z0+1 ..
//      z1-2 its moments go to aux_out in the layout of k_moments.
// Rows 0 and th-1 and planes z0 and z1-1 of each tile are its edge sites:
// some of their sources lie in other work-groups' tiles, and
// k_batch_edge_moments pulls their moments from fout afterwards.
constexpr int CM_NXY = 16;  // directions with an in-plane component
// LDS slot of direction i (-1: c_x = c_y = 0, the site's own populations)
__device__ constexpr int CM_SLOT[19] = {0, 1, 2, 3, -1, -1, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, -1};

__device__ __forceinline__ uint4 cm_item(const uint4 *__restrict__ items, uint32_t it)
{
    uint4 w = items[it];
    w.x = __builtin_amdgcn_readfirstlane(w.x);
    w.y = __builtin_amdgcn_readfirstlane(w.y);
    w.z = __builtin_amdgcn_readfirstlane(w.z);
    w.w = __builtin_amdgcn_readfirstlane(w.w);
    return w;
}

// The moments of one site (psi, rho per component, the dipole-relaxation
// 4-vector) to the aux field, as moments_site writes them.
template <typename R, bool Dev>
__device__ __forceinline__ void put_aux3(const Phys<R> &P, R *__restrict__ aux, uint32_t s, const R (&rho)[3],
                                         R da0, R da1, R da2)
{
#pragma unroll
    for (int c = 0; c < 3; c++) {
        st_plane(aux + (size_t)c * P.ns, s, psi_of<Dev>(rho[c], P.emr0[c]));
        st_plane(aux + (size_t)(3 + c) * P.ns, s, rho[c]);
    }
    V4<R> *vec = (V4<R> *)(aux + aux3_vec_off(P.ns));
    vec[s] = V4<R>{rho[1] - rho[0], da0, da1, da2};
}

template <typename R, bool Dev, int BD, int NTHR>
__global__ __launch_bounds__(NTHR) __attribute__((amdgpu_waves_per_eu(4))) void k_batch_cm(
    const BatchLat<R> *__restrict__ lat, const uint4 *__restrict__ items, uint32_t nitems, const R *__restrict__ f,
    const R *__restrict__ aux, const R *__restrict__ db, R *__restrict__ fout, R *__restrict__ aux_out, uint32_t xg)
{
    static_assert(Dev, "the collide-moments plan runs the FP32 deviation form (ZORD order)");
    // post-collision populations of plane k, two components at a time (the
    // red and blue ones, then the surfactant's): 140 B per thread, so that
    // four work-groups of 256 threads (two of 512) share a CU
    __shared__ R xb[2][CM_NXY][NTHR];
    __shared__ R dbs[3][NTHR];  // dipole field of plane k
    const int tid = threadIdx.x;
    const uint32_t n8 = xcd_round(nitems);
    for (uint32_t t = blockIdx.x; t < n8; t += gridDim.x) {
        const uint32_t it = xcd_tile(t, n8, xg);
        if (it >= nitems) continue;  // uniform over the block
        const uint4 w = cm_item(items, it);
        const BatchLat<R> &L = lat[w.x];
        if (L.P.bdist != BD) continue;  // uniform: this launch runs one equilibrium form
        const Phys<R> &P = L.P;
        const R *fl = f + L.f_off;
        R *fo = fout + L.f_off;
        const R *al = aux + L.aux_off;
        R *ao = aux_out + L.aux_off;
        const R *dl = db + L.db_off;
        const int x0 = (int)(w.y & 0xffffu), y0 = (int)(w.y >> 16);
        const int z0 = (int)(w.z & 0xffffu), z1 = (int)(w.z >> 16);
        const int tw = (int)(w.w & 511u), th = (int)((w.w >> 9) & 511u), S = (int)(w.w >> 18);
        const int rx = tid % S, ry = tid / S;
        const bool act = ry < th && rx <= tw + 1;
        const bool own = act && rx >= 1 && rx <= tw;
        const bool gat = own && ry >= 1 && ry <= th - 2;
        const int x = wrap1(x0 - 1 + rx, P.nx), y = act ? y0 + ry : 0;
        // densities (component, plane k-1 / k / k+1) and advected dipole
        // (plane, axis) of the own site, accumulated over the source planes
        R ar[3][3], ad[3][3];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) ar[a][b] = ad[a][b] = R(0);
        // one component's gather from plane k: LDS buffer u for the
        // directions with an in-plane component, the site's own registers
        // for the others
        auto gather = [&](int c, int u, const R (&own_z)[3], const R (&dbo)[3]) __attribute__((always_inline)) {
            int tq = tid, sq = S;  // opaque: the source slots are computed here, not hoisted
            asm volatile("" : "+v"(tq), "+s"(sq));
#pragma unroll
            for (int ii = 0; ii < 19; ii++) {
                const int i = ZORD[ii];
                const int q = tq - CX[i] - CY[i] * sq;  // source region slot
                const R v = CM_SLOT[i] < 0 ? own_z[i == 4 ? 0 : (i == 5 ? 1 : 2)] : xb[u][CM_SLOT[i] < 0 ? 0 : CM_SLOT[i]][q];
                const int d = CZ[i] + 1;  // destination plane k - 1, k or k + 1
                ar[c][d] = ar[c][d] + v * R(GD[i]);
                if (c == 2) {
                    const R ns_ = Dev ? (i < 18 ? P.wrho0_move[2] : P.wrho0_rest[2]) + v : v;
                    const R ng = ns_ * R(GD[i]);
                    const R b0 = CM_SLOT[i] < 0 ? dbo[0] : dbs[0][q];
                    const R b1 = CM_SLOT[i] < 0 ? dbo[1] : dbs[1][q];
                    const R b2 = CM_SLOT[i] < 0 ? dbo[2] : dbs[2][q];
                    da_term<Dev>(ng, b0, b1, b2, ad[d][0], ad[d][1], ad[d][2]);
                }
            }
        };
        for (int k = z0; k < z1; k++) {
            // the site's own populations of the directions without an
            // in-plane component (4, 5 and the rest direction 18), and the
            // surfactant's others until they go to LDS
            R nz[3][3], n2[CM_NXY];
            R dbx[3] = {R(0), R(0), R(0)};
            __syncthreads();  // every gather of plane k-1 is done with xb and dbs
            if (act) {
                // opaque copies of the lane coordinates: keeps the compiler
                // from hoisting the per-direction neighbour offsets out of
                // the z loop (dozens of registers held across the march)
                int xo = x, yo = y;
                asm volatile("" : "+v"(xo), "+v"(yo));
                const uint32_t s = sidx(xo, yo, k, P.nx, P.ny);
                R F[3][3];
                sc_forces<R, Dev, 3, false>(P, al, nullptr, dl, xo, yo, k, s, F);
                asm volatile("" ::: "memory");
                R n[3][19];
#pragma unroll
                for (int i = 0; i < 19; i++)
#pragma unroll
                    for (int c = 0; c < 3; c++) n[c][i] = pull<R, 3, false>(fl, nullptr, P, c, i, xo, yo, k);
                R rho[3], mom[3][3];
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    rho[c] = R(0);
                    mom[c][0] = mom[c][1] = mom[c][2] = R(0);
                }
#pragma unroll
                for (int ii = 0; ii < 19; ii++) {
                    const int i = ZORD[ii];
#pragma unroll
                    for (int c = 0; c < 3; c++) acc_moments<R>(i, n[c][i], rho[c], mom[c]);
                }
                // plain stores: non-temporal ones measured 6-26% slower here
                // (tile row segments narrower than, or straddling, cache lines)
                bgk_relax<R, Dev, 3, BD, false>(
                    P, [&](int c, int i) __attribute__((always_inline)) { return n[c][i]; }, rho, mom, F, xo, yo, k,
                    s, fo, own, [&](int c, int i, R nv) __attribute__((always_inline)) {
                        if (CM_SLOT[i] < 0)
                            nz[c][i == 4 ? 0 : (i == 5 ? 1 : 2)] = nv;
                        else if (c < 2)
                            xb[c][CM_SLOT[i] < 0 ? 0 : CM_SLOT[i]][tid] = nv;
                        else
                            n2[CM_SLOT[i] < 0 ? 0 : CM_SLOT[i]] = nv;
                    });
                dbx[0] = ld_plane(dl, s);
                dbx[1] = ld_plane(plane_of(dl, 1u, P.ns), s);
                dbx[2] = ld_plane(plane_of(dl, 2u, P.ns), s);
#pragma unroll
                for (int a = 0; a < 3; a++) dbs[a][tid] = dbx[a];
            }
            __syncthreads();
            if (gat) {
                gather(0, 0, nz[0], dbx);
                gather(1, 1, nz[1], dbx);
            }
            __syncthreads();  // the red populations' buffer is free
            if (act) {
#pragma unroll
                for (int j = 0; j < CM_NXY; j++) xb[0][j][tid] = n2[j];
            }
            __syncthreads();
            if (gat) {
                gather(2, 0, nz[2], dbx);
                // plane k-1 has received all its directions
                if (k - 1 >= z0 + 1 && k - 1 <= z1 - 2) {
                    const R rho[3] = {ar[0][0], ar[1][0], ar[2][0]};
                    put_aux3<R, Dev>(P, ao, sidx(x, y, k - 1, P.nx, P.ny), rho, ad[0][0], ad[0][1], ad[0][2]);
                }
            }
#pragma unroll
            for (int a = 0; a < 3; a++) {
                ar[a][0] = ar[a][1];
                ar[a][1] = ar[a][2];
                ar[a][2] = R(0);
                ad[0][a] = ad[1][a];
                ad[1][a] = ad[2][a];
                ad[2][a] = R(0);
            }
        }
    }
}

// The moments of the edge sites of k_batch_cm's tiles (the aux layout of
// k_moments), pulled from the post-collision populations f.  Work item:
// (lattice, first entry of its run of 256 site indices); ~0u entries pad a
// lattice's list to whole runs.
template <typename R, bool Dev>
__global__ __launch_bounds__(256) void k_batch_edge_moments(const BatchLat<R> *__restrict__ lat,
                                                            const uint2 *__restrict__ items, uint32_t nitems,
                                                            const uint32_t *__restrict__ sites,
                                                            const R *__restrict__ f, const R *__restrict__ db,
                                                            R *__restrict__ aux, uint32_t xg)
{
    const uint32_t n8 = xcd_round(nitems);
    for (uint32_t t = blockIdx.x; t < n8; t += gridDim.x) {
        uint32_t it = xcd_tile(t, n8, xg);
        if (it >= nitems) continue;
        uint2 w = batch_item(items, it);
        const BatchLat<R> &L = lat[w.x];
        const uint32_t s = sites[w.y + threadIdx.x];
        if (s < L.P.ns) moments_site<R, Dev, 3, false>(L.P, f + L.f_off, nullptr, db + L.db_off, aux + L.aux_off, s);
    }
}

static __global__ void k_le_neighbours(LEState *L, int nz, int nt_set)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    *L = le_advance(*L, nz, nt_set);
}

}  // namespace lb3d
