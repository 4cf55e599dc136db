// lb3d_solver.h -- host-side interface between the C ABI (lb3d_gpu.hip) and
// the precision-specific kernel launchers (lb3d_fp64.hip, lb3d_fp32.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/lb3d_gpu.h"
#include "lb3d_lattice.h"

namespace lb3d {

// Device buffers of one lattice.  f[cur] holds the post-collision populations
// of step nt; collide writes f[1-cur].
struct DevBufs {
    void *f[2] = {nullptr, nullptr};
    // ternary: db is the dipole field of the state (read by the moments and
    // the one-pass kernel), db_alt receives the relaxed field of the step
    // (written by the dipole pass or the one-pass kernel, read by the collide);
    // a single lattice alternates the two with the population parity, the
    // batch relaxes one field in place (db_alt == db)
    void *aux = nullptr, *db = nullptr, *face_pc = nullptr, *face_aux = nullptr;
    void *db_alt = nullptr;
    // Lees-Edwards edge columns (x = nx-1 and x = 0) in a compact layout: the
    // post-collision populations (edge_pc) and the streamed ones (edge_adv)
    void *edge_pc = nullptr, *edge_adv = nullptr;
    const uint8_t *rock = nullptr;  // 1 at rock sites, or nullptr for a lattice without walls
    LEState *le = nullptr;
    int cur = 0;
};

// One precision's kernels for one parameter set.
class Launcher {
  public:
    virtual ~Launcher() {}
    // chunk of consecutive tiles per XCD in the stepping kernels (xcd_tile);
    // 1 is the plain round-robin order
    uint32_t xg = 1;
    bool walls = false;  // the lattice has rock sites (the one-pass plans are then unavailable)
    virtual void moments(const DevBufs &b, hipStream_t s) = 0;
    virtual void dipole(const DevBufs &b, hipStream_t s) = 0;
    virtual void collide(const DevBufs &b, hipStream_t s) = 0;  // f[cur] -> f[1-cur]
    // Lees-Edwards x-halo planes, pre- (post = false) or post-advection;
    // compact: from the edge columns (b.edge_pc written by the two-pass
    // collide or le_gather_edges, b.edge_adv by the moments pass)
    // adv: LE_STATE_AS_IS, or the halo-plane kernels compute the step's state
    // from the stored one (le_advance with nt_set = adv) and the next collide
    // stores it (set_le_commit)
    virtual void le_faces(const DevBufs &b, bool post, bool compact, hipStream_t s, int adv = LE_STATE_AS_IS) = 0;
    virtual void set_le_commit(int adv) = 0;
    // b.edge_pc from the populations f[cur]
    virtual void le_gather_edges(const DevBufs &b, hipStream_t s) = 0;
    // the Shan-Chen force field from b.aux and b.db into out[(c*3 + k)*ns + s]
    // (working type R), periodic in x
    virtual void forces(const DevBufs &b, void *out, hipStream_t s) = 0;
    // the whole binary step in one kernel (f[cur] -> f[1-cur]); false if the
    // tile shape has no instantiation or the form is not binary
    // stage: where plane k's populations wait for the collide: 0 registers,
    // 1 LDS (both pulled one plane ahead), 2 pulled again by the collide
    virtual bool fused(const DevBufs &b, int tx, int ty, int lz, int stage, hipStream_t s) = 0;
    virtual bool fused_supported(int tx, int ty, int stage) = 0;
    virtual void init_noise(const DevBufs &b, uint64_t seed, const double rho_mean[3],
                            const double amp[3], int dipoles, hipStream_t s) = 0;
    // per-block FP64 partial sums for the sanity reductions (NDIAG per block)
    virtual void diagnostics(const DevBufs &b, double *partials, hipStream_t s) = 0;
    // per-block partials of the full sanity check (NSAN per block)
    virtual void sanity(const DevBufs &b, double *partials, hipStream_t s) = 0;
    virtual unsigned diag_blocks() const = 0;
    virtual size_t real_bytes() const = 0;
};

// Tile shapes of k_fused2 that are instantiated (TX, TY).
constexpr int FUSED_SHAPES[][2] = {{64, 4}, {32, 8}, {64, 8}, {32, 16}, {128, 4}, {32, 4}, {16, 16},
                                   {128, 2}, {256, 2}};
constexpr int N_FUSED_SHAPES = sizeof(FUSED_SHAPES) / sizeof(FUSED_SHAPES[0]);

constexpr int NDIAG = 9;  // mass[3], momentum[3], nan sites, negative sites, fast sites
// full sanity partials: NDIAG values, fluid sites, velocity sums [3][3], then
// per (species, axis) the largest |velocity| and its loop-order key
constexpr int SAN_FLUID = 9, SAN_VSUM = 10, SAN_MAX = 19, SAN_KEY = 28, NSAN = 37;
// fill *out from nb blocks of k_sanity partials (host)
void reduce_sanity(const double *partials, unsigned nb, int ncomp, int nx, int ny, int nz, lb3d_sanity *out);
constexpr unsigned BLOCK = 256;

// Batched ensemble kernels of one precision.  The lattice table (BatchLat<R>
// per lattice) and the work list live in device memory.
class BatchLauncher {
  public:
    virtual ~BatchLauncher() {}
    uint32_t xg = 0;  // chunk of consecutive work items per XCD (xcd_tile)
    virtual size_t lat_bytes() const = 0;   // sizeof(BatchLat<R>)
    virtual void set_bdist_mask(int mask) = 0;
    // bytes of one copy of the batch's populations (chooses the collide's store form)
    virtual void set_pop_bytes(uint64_t bytes) = 0;
    // host image of the lattice table entry k
    virtual void fill_lat(void *dst, const lb3d_params &p, uint64_t f_off, uint64_t aux_off,
                          uint64_t db_off) const = 0;
    // phase ph (0 moments, 1 dipole, 2 collide) over nitems work items;
    // packed: the ternary moments and dipole phases use the packed layout of
    // the moments field (one psi_r, psi_b, psi_s, rho_s 4-vector per site),
    // else the planes; the ternary collide phase always reads the packed one
    virtual void phase(int ph, const void *lat, const void *items, uint32_t nitems, unsigned nwg,
                       const void *f, void *fout, void *aux, void *db, int ncomp, hipStream_t s,
                       bool packed) = 0;
    // the collide-moments plan (FP32 ternary only): k_batch_cm over nitems
    // tiles (uint4 items) with work-groups of threads = 256 or 512, and k_batch_edge_moments over nitems runs of 256
    // edge sites (uint2 items into the site list)
    virtual bool cm_supported(int ncomp) const = 0;
    virtual void cm_collide(const void *lat, const void *items, uint32_t nitems, unsigned nwg, int threads,
                            const void *f, void *fout, const void *aux_in, void *aux_out, const void *db,
                            hipStream_t s) = 0;
    virtual void edge_moments(const void *lat, const void *items, uint32_t nitems, unsigned nwg, const void *sites,
                              const void *f, const void *db, void *aux, hipStream_t s) = 0;
};

Launcher *make_launcher_fp64(const lb3d_params &p);
// one two-pass timestep over caller-provided buffers (no allocation)
void step_buffers_fp64(const lb3d_params &p, const DevBufs &b, hipStream_t s);
void step_buffers_fp32(const lb3d_params &p, const DevBufs &b, hipStream_t s);
Launcher *make_launcher_fp32(const lb3d_params &p);
BatchLauncher *make_batch_launcher_fp64();
BatchLauncher *make_batch_launcher_fp32();

// Error state of the C ABI (thread-local message, returns code).
int set_error(int code, const char *msg);
void clear_error();
// Host conversion of the reference's AoS lattice (lbe_site records) into the
// device layout of one lattice: populations (FP64, or FP32 deviations from
// w_i*rho0) and the ternary dipole field.  Returns 0 or an LB3D_E* code.
int aos_to_device_layout(const lb3d_params &p, const double *whole_N, int halo, int site_doubles,
                         std::vector<unsigned char> &f_bytes, std::vector<unsigned char> &db_bytes,
                         std::vector<uint8_t> *rock = nullptr, size_t *nrock = nullptr);
// The inverse for the populations: device bytes -> FP64 populations.
void device_layout_to_pops(const lb3d_params &p, const unsigned char *f_bytes, double *out);
double rho0_of(const lb3d_params &p, int c);

// Parameters in the kernels' working type.
template <typename R>
Phys<R> make_phys(const lb3d_params &p);

}  // namespace lb3d
