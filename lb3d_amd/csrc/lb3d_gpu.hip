// lb3d_gpu.hip -- the C ABI of include/lb3d_gpu.h: contexts, device memory,
// the reference's stage sequence, HIP-graph replay, diagnostics and the
// conversion between the reference's AoS lattice and the device layout.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <utility>
#include <string>
#include <vector>

#include "lb3d_kernels.h"
#include "lb3d_solver.h"

using namespace lb3d;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                     \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(LB3D_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
    } while (0)

enum StageState { IDLE = 0, AFTER_BC6 = 1, AFTER_ADV = 2, AFTER_BC8 = 3 };

// HIP events that are destroyed when the scope ends, also on an early error return
struct Events {
    std::vector<hipEvent_t> e;
    explicit Events(size_t n) : e(n, nullptr) {}
    Events(const Events &) = delete;
    Events &operator=(const Events &) = delete;
    ~Events()
    {
        for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
    }
    hipError_t create()
    {
        for (hipEvent_t &x : e) {
            hipError_t r = hipEventCreate(&x);
            if (r != hipSuccess) return r;
        }
        return hipSuccess;
    }
    hipEvent_t &operator[](size_t i) { return e[i]; }
};

const char *const KNAMES[] = {"le_neighbours", "le_faces_pc", "moments",
                              "le_faces_aux",  "dipole",      "collide", "fused"};
enum { K_LEN = 0, K_LE6, K_MOM, K_LE8, K_DIP, K_COL, K_FUS, K_COUNT };

// How a timestep is executed: the two-pass sequence (moments, [dipole],
// collide) or, for binary lattices, the one-pass z-march kernel with a tile
// shape (tx, ty) and a z-chunk lz.
const char *const STAGE_NAMES[3] = {"regs", "lds", "reload"};

struct Plan {
    bool fused = false;
    int tx = 64, ty = 4, lz = 0;
    int stage = 1;      // plane k's populations: 0 registers, 1 LDS, 2 pulled again
    uint32_t xg = 1;    // tiles per XCD chunk (xcd_tile): 1 plain order, 0 one run per XCD
};

}  // namespace

struct lb3d_ctx {
    lb3d_params p{};
    Launcher *L = nullptr;
    DevBufs b;
    hipStream_t stream = nullptr;
    int nt = 0;
    int state = IDLE;
    bool le = false;
    bool aux_valid = false;
    LEState host_le{};
    hipGraphExec_t graph = nullptr;
    double *d_partials = nullptr;
    uint8_t *rock_dev = nullptr;  // rock mask (lattices with walls)
    size_t f_bytes = 0, aux_bytes = 0, db_bytes = 0;
    size_t aux_elems = 0;  // elements of the aux field
    Plan plan;
    bool tuned = false;
    bool aux_from_prev = false;  // aux is recomputable from f[1-cur] (fused binary steps)
    // ternary dipole fields, tied to the population parity: a step reads
    // dbk[cur] (the field of the state) and writes dbk[1-cur] (the relaxed
    // field), so a captured two-step graph finds the live field where it
    // expects it whatever the host did before
    void *dbk[2] = {nullptr, nullptr};
    // Lees-Edwards: the edge columns b.edge_pc hold f[cur]'s (written by the
    // two-pass collide, else gathered before the next halo-plane pass)
    bool edges_ok = false;
};

// The halo-plane kernels read the compact edge columns on the two-pass
// plan, whose moments and collide passes write them.
static bool compact_edges(const lb3d_ctx *c) { return !c->plan.fused; }

static void ensure_edges(lb3d_ctx *c)
{
    if (c->le && compact_edges(c) && !c->edges_ok) {
        c->L->le_gather_edges(c->b, c->stream);
        c->edges_ok = true;
    }
}

// Points b.db / b.db_alt at the dipole fields of the current parity.
static void set_cur(lb3d_ctx *c, int cur)
{
    c->b.cur = cur;
    c->b.db = c->dbk[cur];
    c->b.db_alt = c->dbk[1 - cur];
}

static void flip(lb3d_ctx *c) { set_cur(c, 1 - c->b.cur); }

static void drop_graph(lb3d_ctx *c)
{
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
}

static int check_params(const lb3d_params *p)
{
    if (!p) return fail(LB3D_EINVAL, "params is NULL");
    if (p->nx < 2 || p->ny < 1 || p->nz < 1)
        return fail(LB3D_EINVAL, "lattice extent must be nx>=2, ny>=1, nz>=1");
    if (p->ncomp != 2 && p->ncomp != 3) return fail(LB3D_EINVAL, "ncomp must be 2 or 3");
    if (p->bdist < 0 || p->bdist > 6)
        return fail(LB3D_EINVAL, "bdist must be 0-6 (code/lb3d_bdist.F90; 7 computes nothing there)");
    if (p->precision != LB3D_FP64 && p->precision != LB3D_FP32_DEV)
        return fail(LB3D_EINVAL, "precision must be LB3D_FP64 or LB3D_FP32_DEV");
    if (p->boundary != LB3D_BC_PERIODIC && p->boundary != LB3D_BC_LEES_EDWARDS)
        return fail(LB3D_EINVAL, "boundary must be periodic or Lees-Edwards");
    if (p->boundary == LB3D_BC_LEES_EDWARDS && p->ncomp == 3)
        // code/lb3d_compatibility.F90:105-108
        return fail(LB3D_EINVAL, "Lees-Edwards boundaries are not supported with surfactant");
    for (int c = 0; c < p->ncomp; c++)
        if (!(p->tau[c] > 0.0) || !(p->amass[c] > 0.0))
            return fail(LB3D_EINVAL, "tau and amass must be positive");
    if (p->ncomp == 3 && !(p->tau_d > 0.0)) return fail(LB3D_EINVAL, "tau_d must be positive");
    if (p->inv_fluid != 0 && p->inv_fluid != 12)
        return fail(LB3D_EINVAL, "inv_fluid must be 0 (no invasion) or 12 (the Zou-He flux boundary)");
    if (p->inv_fluid == 12 && p->boundary == LB3D_BC_LEES_EDWARDS)
        return fail(LB3D_EINVAL, "the Zou-He boundary needs a periodic lattice");
    if (p->inv_fluid == 12 && (p->inv_type < -1 || p->inv_type > 1))
        return fail(LB3D_EINVAL, "inv_type must be -1, 0 or 1 (code/lb3d_invasion.F90:2136-2140)");
    double sites = (double)p->nx * (double)p->ny * (double)p->nz;
    if (sites * 19.0 * p->ncomp >= 4294967296.0)
        return fail(LB3D_EINVAL, "lattice too large for 32-bit population indexing");
    return LB3D_OK;
}

namespace lb3d {
void reduce_sanity(const double *h, unsigned nb, int ncomp, int nx, int ny, int nz, lb3d_sanity *out)
{
    std::memset(out, 0, sizeof *out);
    double acc[SAN_MAX] = {0};
    double mx[9], key[9];
    for (int k = 0; k < 9; k++) mx[k] = key[k] = -1.0;
    for (unsigned b = 0; b < nb; b++) {
        const double *p = h + (size_t)b * NSAN;
        for (int k = 0; k < SAN_MAX; k++) acc[k] = acc[k] + p[k];
        for (int k = 0; k < 9; k++) {
            double m = p[SAN_MAX + k], q = p[SAN_KEY + k];
            if (m > mx[k] || (m == mx[k] && q > key[k])) {
                mx[k] = m;
                key[k] = q;
            }
        }
    }
    for (int k = 0; k < 3; k++) {
        out->mass[k] = k < ncomp ? acc[k] : 0.0;
        out->momentum[k] = acc[3 + k];
    }
    out->nan_sites = (int64_t)llround(acc[6]);
    out->negative_sites = (int64_t)llround(acc[7]);
    out->fast_sites = (int64_t)llround(acc[8]);
    out->fluid_sites = (int64_t)llround(acc[SAN_FLUID]);
    for (int c = 0; c < 3; c++)
        for (int a = 0; a < 3; a++) {
            int k = 3 * c + a;
            out->vel_sum[c][a] = c < ncomp ? acc[SAN_VSUM + k] : 0.0;
            if (c < ncomp && key[k] >= 0.0) {
                long long q = llround(key[k]);
                out->max_vel[c][a] = mx[k];
                out->max_pos[c][a][0] = (int32_t)(q / ((long long)ny * nz)) + 1;
                out->max_pos[c][a][1] = (int32_t)((q / nz) % ny) + 1;
                out->max_pos[c][a][2] = (int32_t)(q % nz) + 1;
            }
        }
    (void)nx;
}

int set_error(int code, const char *msg) { return fail(code, msg); }
void clear_error() { g_err.clear(); }
double rho0_of(const lb3d_params &p, int k) { return p.rho0[k] > 0.0 ? p.rho0[k] : 0.7; }

int aos_to_device_layout(const lb3d_params &p, const double *N, int h, int sd,
                         std::vector<unsigned char> &f_bytes, std::vector<unsigned char> &db_bytes,
                         std::vector<uint8_t> *rock, size_t *nrock)
{
    const int C = p.ncomp;
    if (sd != (C == 3 ? 79 : 49)) return fail(LB3D_EINVAL, "site_doubles must be 79 (ternary) or 49 (binary)");
    if (h < 1) return fail(LB3D_EINVAL, "halo_extent must be >= 1");
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const size_t ns = (size_t)nx * ny * nz;
    const size_t lx = nx + 2 * h, ly = ny + 2 * h;
    const int rock_off = C == 3 ? 78 : 48;
    const bool dev = p.precision == LB3D_FP32_DEV;
    const size_t rs = dev ? 4 : 8;
    f_bytes.assign((size_t)C * 19 * ns * rs, 0);
    db_bytes.assign(C == 3 ? 3 * ns * rs : 0, 0);
    float *ff = (float *)f_bytes.data();
    double *fd = (double *)f_bytes.data();
    float *dbf = (float *)db_bytes.data();
    double *dbd = (double *)db_bytes.data();
    size_t nr = 0;
    if (rock) rock->assign(ns, 0);
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const double *S = N + ((size_t)(x + h) + lx * ((size_t)(y + h) + ly * (size_t)(z + h))) * sd;
                size_t s = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z);
                if (S[rock_off] != 0.0) {
                    if (!rock) return fail(LB3D_EINVAL, "rock sites are not supported here (rock_state must be 0)");
                    (*rock)[s] = 1;
                    nr++;
                }
                for (int k = 0; k < C; k++)
                    for (int i = 0; i < 19; i++) {
                        double v = S[24 * k + i];
                        size_t q = ((size_t)k * 19 + i) * ns + s;
                        if (dev)
                            ff[q] = (float)(v - (i < 18 ? 1.0 / 36.0 : 1.0 / 3.0) * rho0_of(p, k));
                        else
                            fd[q] = v;
                    }
                if (C == 3)
                    for (int k = 0; k < 3; k++) {
                        if (dev)
                            dbf[k * ns + s] = (float)S[75 + k];
                        else
                            dbd[k * ns + s] = S[75 + k];
                    }
            }
    if (nrock) *nrock = nr;
    return LB3D_OK;
}

void device_layout_to_pops(const lb3d_params &p, const unsigned char *fb, double *out)
{
    const int C = p.ncomp;
    const size_t ns = (size_t)p.nx * p.ny * p.nz;
    if (p.precision == LB3D_FP32_DEV) {
        const float *h = (const float *)fb;
        for (int k = 0; k < C; k++)
            for (int i = 0; i < 19; i++) {
                double base = (i < 18 ? 1.0 / 36.0 : 1.0 / 3.0) * rho0_of(p, k);
                size_t off = ((size_t)k * 19 + i) * ns;
                for (size_t q = 0; q < ns; q++) out[off + q] = base + (double)h[off + q];
            }
    } else {
        std::memcpy(out, fb, (size_t)C * 19 * ns * sizeof(double));
    }
}
}  // namespace lb3d

extern "C" {

int lb3d_gpu_abi_version(void) { return LB3D_ABI_VERSION; }

const char *lb3d_gpu_last_error(void) { return g_err.c_str(); }

int lb3d_gpu_create(const lb3d_params *p, lb3d_ctx **out)
{
    g_err.clear();
    if (!out) return fail(LB3D_EINVAL, "out is NULL");
    *out = nullptr;
    int rc = check_params(p);
    if (rc) return rc;
    lb3d_ctx *c = new lb3d_ctx();
    c->p = *p;
    c->le = p->boundary == LB3D_BC_LEES_EDWARDS;
    c->L = p->precision == LB3D_FP64 ? make_launcher_fp64(*p) : make_launcher_fp32(*p);
    size_t rs = c->L->real_bytes();
    size_t ns = (size_t)p->nx * p->ny * p->nz;
    c->aux_elems = aux_elems(p->ncomp, ns);
    c->f_bytes = (size_t)p->ncomp * 19 * ns * rs;
    c->aux_bytes = c->aux_elems * rs;
    c->db_bytes = p->ncomp == 3 ? 3 * ns * rs : 0;
    size_t plane = (size_t)p->ny * p->nz;
    auto cleanup = [&](int code, const std::string &m) {
        lb3d_gpu_destroy(c);
        return fail(code, m);
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return cleanup(LB3D_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    for (int k = 0; k < 2; k++)
        if ((e = hipMalloc(&c->b.f[k], c->f_bytes)) != hipSuccess)
            return cleanup(LB3D_ENOMEM, std::string("hipMalloc populations: ") + hipGetErrorString(e));
    if ((e = hipMalloc(&c->b.aux, c->aux_bytes)) != hipSuccess)
        return cleanup(LB3D_ENOMEM, std::string("hipMalloc aux: ") + hipGetErrorString(e));
    if (c->db_bytes && ((e = hipMalloc(&c->dbk[0], c->db_bytes)) != hipSuccess ||
                        (e = hipMalloc(&c->dbk[1], c->db_bytes)) != hipSuccess))
        return cleanup(LB3D_ENOMEM, std::string("hipMalloc db: ") + hipGetErrorString(e));
    set_cur(c, 0);
    if (c->le) {
        if ((e = hipMalloc(&c->b.edge_pc, 2 * 2 * 19 * plane * rs)) != hipSuccess ||
            (e = hipMalloc(&c->b.edge_adv, 2 * 2 * 19 * plane * rs)) != hipSuccess)
            return cleanup(LB3D_ENOMEM, std::string("hipMalloc LE edge columns: ") + hipGetErrorString(e));
        if ((e = hipMalloc(&c->b.face_pc, 2 * 2 * 19 * plane * rs)) != hipSuccess ||
            (e = hipMalloc(&c->b.face_aux, 2 * 4 * plane * rs)) != hipSuccess)
            return cleanup(LB3D_ENOMEM, std::string("hipMalloc LE planes: ") + hipGetErrorString(e));
    }
    if ((e = hipMalloc((void **)&c->b.le, sizeof(LEState))) != hipSuccess ||
        (e = hipMalloc((void **)&c->d_partials, (size_t)c->L->diag_blocks() * NSAN * sizeof(double))) !=
            hipSuccess)
        return cleanup(LB3D_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    if (p->ncomp == 2 && c->L->fused_supported(32, 4, true)) {
        // default one-pass plan where it exists; lb3d_gpu_autotune picks the measured best
        c->plan.fused = true;
        c->plan.tx = c->L->fused_supported(64, 4, true) ? 64 : 32;
        c->plan.ty = 4;
        c->plan.lz = p->nz < 64 ? p->nz : 64;
        c->plan.xg = 0;
    }
    c->L->xg = c->plan.xg;
    c->host_le.shear_u = p->shear_u;
    c->host_le.shear_tmp = p->shear_u;
    c->host_le.shear_sum = p->shear_sum;
    c->host_le.shear_omega = p->shear_omega;
    *out = c;
    return LB3D_OK;
}

int lb3d_gpu_destroy(lb3d_ctx *c)
{
    if (!c) return LB3D_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    for (int k = 0; k < 2; k++)
        if (c->b.f[k]) (void)hipFree(c->b.f[k]);
    if (c->b.aux) (void)hipFree(c->b.aux);
    for (int k = 0; k < 2; k++)
        if (c->dbk[k]) (void)hipFree(c->dbk[k]);
    if (c->b.edge_pc) (void)hipFree(c->b.edge_pc);
    if (c->b.edge_adv) (void)hipFree(c->b.edge_adv);
    if (c->b.face_pc) (void)hipFree(c->b.face_pc);
    if (c->b.face_aux) (void)hipFree(c->b.face_aux);
    if (c->b.le) (void)hipFree(c->b.le);
    if (c->d_partials) (void)hipFree(c->d_partials);
    if (c->rock_dev) (void)hipFree(c->rock_dev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c->L;
    delete c;
    return LB3D_OK;
}

static int push_le(lb3d_ctx *c)
{
    HIP_TRY(hipMemcpyAsync(c->b.le, &c->host_le, sizeof(LEState), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LB3D_OK;
}

static int pull_le(lb3d_ctx *c)
{
    HIP_TRY(hipMemcpyAsync(&c->host_le, c->b.le, sizeof(LEState), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LB3D_OK;
}

static bool is_dev(const lb3d_ctx *c) { return c->p.precision == LB3D_FP32_DEV; }
static double rho0_of(const lb3d_ctx *c, int k) { return lb3d::rho0_of(c->p, k); }
static double w_of(int i) { return i < 18 ? 1.0 / 36.0 : 1.0 / 3.0; }

// Rock mask on the device (nullptr: no walls).  With walls only the two-pass
// plan exists, so a one-pass plan falls back to it.
static int set_walls(lb3d_ctx *c, const uint8_t *mask)
{
    const size_t ns = (size_t)c->p.nx * c->p.ny * c->p.nz;
    // a captured graph holds the kernels of the old wall state
    drop_graph(c);
    if (!mask) {
        c->L->walls = false;
        c->b.rock = nullptr;
        return LB3D_OK;
    }
    if (!c->rock_dev) HIP_TRY(hipMalloc((void **)&c->rock_dev, ns));
    HIP_TRY(hipMemcpyAsync(c->rock_dev, mask, ns, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->b.rock = c->rock_dev;
    c->L->walls = true;
    if (c->plan.fused) {
        c->plan = Plan();
        c->L->xg = c->plan.xg;
    }
    return LB3D_OK;
}

int lb3d_gpu_upload_sites(lb3d_ctx *c, const double *N, int32_t h, int32_t sd, int32_t nt0)
{
    g_err.clear();
    if (!c || !N) return fail(LB3D_EINVAL, "NULL argument");
    std::vector<unsigned char> fb, dbb;
    std::vector<uint8_t> rock;
    size_t nrock = 0;
    int rc = aos_to_device_layout(c->p, N, h, sd, fb, dbb, &rock, &nrock);
    if (rc) return rc;
    if (nrock && (c->p.ncomp != 2 || c->le))
        return fail(LB3D_EINVAL, "rock sites (walls) are supported on periodic binary lattices");
    if (nrock && c->p.inv_fluid == 12)
        return fail(LB3D_EINVAL, "rock sites (walls) and the Zou-He boundary are not supported together");
    if ((rc = set_walls(c, nrock ? rock.data() : nullptr))) return rc;
    set_cur(c, 0);
    HIP_TRY(hipMemcpyAsync(c->b.f[0], fb.data(), c->f_bytes, hipMemcpyHostToDevice, c->stream));
    if (c->p.ncomp == 3)
        HIP_TRY(hipMemcpyAsync(c->b.db, dbb.data(), c->db_bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->nt = nt0;
    c->host_le.nt = nt0;
    c->state = IDLE;
    c->aux_valid = false;
    c->aux_from_prev = false;
    c->edges_ok = false;
    return LB3D_OK;
}

int lb3d_gpu_init_noise(lb3d_ctx *c, uint64_t seed, const double rho_mean[3], const double amp[3],
                        int32_t dipoles)
{
    g_err.clear();
    if (!c || !rho_mean || !amp) return fail(LB3D_EINVAL, "NULL argument");
    int rc = set_walls(c, nullptr);
    if (rc) return rc;
    set_cur(c, 0);
    c->L->init_noise(c->b, seed, rho_mean, amp, dipoles, c->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->nt = 0;
    c->host_le.nt = 0;
    c->state = IDLE;
    c->aux_valid = false;
    c->aux_from_prev = false;
    c->edges_ok = false;
    return LB3D_OK;
}

int lb3d_gpu_get_populations(lb3d_ctx *c, double *out)
{
    g_err.clear();
    if (!c || !out) return fail(LB3D_EINVAL, "NULL argument");
    std::vector<unsigned char> h(c->f_bytes);
    HIP_TRY(hipMemcpyAsync(h.data(), c->b.f[c->b.cur], c->f_bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    device_layout_to_pops(c->p, h.data(), out);
    return LB3D_OK;
}

// The one-pass plans keep no moments: rebuild those of the last advection
// from the previous post-collision populations f[1-cur] and the dipole field
// that step read (db_alt).
static int rebuild_aux(lb3d_ctx *c)
{
    if (c->aux_valid || !c->aux_from_prev) return LB3D_OK;
    DevBufs prev = c->b;
    prev.cur = 1 - c->b.cur;
    if (c->p.ncomp == 3) prev.db = c->b.db_alt;
    c->L->moments(prev, c->stream);
    HIP_TRY(hipGetLastError());
    c->aux_valid = true;
    c->aux_from_prev = false;
    return LB3D_OK;
}

int lb3d_gpu_download_sites(lb3d_ctx *c, double *N, int32_t h, int32_t sd)
{
    g_err.clear();
    if (!c || !N) return fail(LB3D_EINVAL, "NULL argument");
    const int C = c->p.ncomp;
    if (sd != (C == 3 ? 79 : 49)) return fail(LB3D_EINVAL, "site_doubles must be 79 (ternary) or 49 (binary)");
    if (h < 1) return fail(LB3D_EINVAL, "halo_extent must be >= 1");
    const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    const size_t ns = (size_t)nx * ny * nz;
    std::vector<double> pops((size_t)C * 19 * ns);
    int rc = lb3d_gpu_get_populations(c, pops.data());
    if (rc) return rc;
    const bool dev = is_dev(c);
    const size_t rs = c->L->real_bytes();
    std::vector<double> aux, db;
    auto fetch = [&](const void *src, size_t count, std::vector<double> &dst) -> int {
        dst.resize(count);
        if (rs == 8) {
            HIP_TRY(hipMemcpyAsync(dst.data(), src, count * 8, hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));
        } else {
            std::vector<float> t(count);
            HIP_TRY(hipMemcpyAsync(t.data(), src, count * 4, hipMemcpyDeviceToHost, c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));
            for (size_t q = 0; q < count; q++) dst[q] = (double)t[q];
        }
        return LB3D_OK;
    };
    if ((rc = rebuild_aux(c))) return rc;
    if (c->aux_valid && (rc = fetch(c->b.aux, c->aux_elems, aux))) return rc;
    if (C == 3 && (rc = fetch(c->b.db, 3 * ns, db))) return rc;
    std::vector<uint8_t> rock;
    if (c->b.rock) {
        rock.resize(ns);
        HIP_TRY(hipMemcpyAsync(rock.data(), c->b.rock, ns, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    const size_t lx = nx + 2 * h, ly = ny + 2 * h, lz = nz + 2 * h;
    const int rock_off = C == 3 ? 78 : 48;
    auto rec = [&](size_t x, size_t y, size_t z) { return N + (x + lx * (y + ly * z)) * sd; };
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                double *S = rec(x + h, y + h, z + h);
                size_t s = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z);
                std::memset(S, 0, sizeof(double) * sd);
                for (int k = 0; k < C; k++) {
                    double rho = 0.0, m[3] = {0.0, 0.0, 0.0};
                    for (int i = 0; i < 19; i++) {
                        double v = pops[((size_t)k * 19 + i) * ns + s];
                        S[24 * k + i] = v;
                        double ng = v * (double)host::GD[i];
                        rho = rho + ng;
                        m[0] = m[0] + ng * (double)host::CX[i];
                        m[1] = m[1] + ng * (double)host::CY[i];
                        m[2] = m[2] + ng * (double)host::CZ[i];
                    }
                    if (c->aux_valid) {
                        double r = aux[(size_t)(C + k) * ns + s];
                        rho = dev ? rho0_of(c, k) + r : r;
                    }
                    S[24 * k + 19] = rho;
                    for (int a = 0; a < 3; a++) S[24 * k + 20 + a] = m[a];
                }
                if (C == 3)
                    for (int a = 0; a < 3; a++) {
                        S[72 + a] = c->aux_valid ? aux[aux3_vec_off(ns) + 4 * s + 1 + a] : 0.0;
                        S[75 + a] = db[(size_t)a * ns + s];
                    }
                S[rock_off] = 0.0;
                if (!rock.empty() && rock[s]) {
                    // rock sites with the colours lbe_init_rock_colour gives them
                    S[rock_off] = 1.0;
                    if (c->p.rock_colour_r > 0) S[23] = c->p.rock_colour_r;
                    if (c->p.rock_colour_b > 0) S[24 + 23] = c->p.rock_colour_b;
                }
            }
    // halo: periodic images of the interior (x, then y over full x, then z)
    for (size_t z = h; z < (size_t)nz + h; z++)
        for (size_t y = h; y < (size_t)ny + h; y++)
            for (int k = 1; k <= h; k++) {
                std::memcpy(rec(h - k, y, z), rec(nx + h - k, y, z), sizeof(double) * sd);
                std::memcpy(rec(nx + h - 1 + k, y, z), rec(h - 1 + k, y, z), sizeof(double) * sd);
            }
    for (size_t z = h; z < (size_t)nz + h; z++)
        for (int k = 1; k <= h; k++)
            for (size_t x = 0; x < lx; x++) {
                std::memcpy(rec(x, h - k, z), rec(x, ny + h - k, z), sizeof(double) * sd);
                std::memcpy(rec(x, ny + h - 1 + k, z), rec(x, h - 1 + k, z), sizeof(double) * sd);
            }
    for (int k = 1; k <= h; k++)
        for (size_t y = 0; y < ly; y++)
            for (size_t x = 0; x < lx; x++) {
                std::memcpy(rec(x, y, h - k), rec(x, y, nz + h - k), sizeof(double) * sd);
                std::memcpy(rec(x, y, nz + h - 1 + k), rec(x, y, h - 1 + k), sizeof(double) * sd);
            }
    (void)lz;
    return LB3D_OK;
}

int lb3d_gpu_bc_init(lb3d_ctx *c)
{
    g_err.clear();
    if (!c) return fail(LB3D_EINVAL, "NULL context");
    if (c->le) {
        // code/lb3d_bc.F90:58-78 then le_neighbours (code/lb3d_bc_leesedwards.F90:86-140)
        LEState &L = c->host_le;
        L.shear_u = ((double)c->p.nx / ((double)c->p.nx - 1.0)) * c->p.shear_u;
        L.shear_tmp = L.shear_u;
        L.shear_sum = c->p.shear_sum;
        L.shear_omega = c->p.shear_omega;
        if (c->p.restore) L.shear_sum = L.shear_sum - L.shear_u;
        L.shear_sum = L.shear_sum + L.shear_u;
        long t = (long)(2.0 * L.shear_sum);
        L.lelo = (int)(t % (long)c->p.nz);
        double lf = 2.0 * L.shear_sum;
        L.lefrac = std::fabs(lf - std::trunc(lf));
        L.nt = c->nt;
        return push_le(c);
    }
    return LB3D_OK;
}

int lb3d_gpu_force_apply(lb3d_ctx *c, int32_t stage)
{
    g_err.clear();
    if (!c) return fail(LB3D_EINVAL, "NULL context");
    // lbe_force_apply: the constant and Kolmogorov forces are outside this path
    if (stage != 6 && stage != 8 && stage != 10) return fail(LB3D_EINVAL, "force_apply stage must be 6, 8 or 10");
    return LB3D_OK;
}

static void launch_le_neighbours(lb3d_ctx *c, int nt_set)
{
    k_le_neighbours<<<1, 64, 0, c->stream>>>(c->b.le, c->p.nz, nt_set);
}

int lb3d_gpu_bc_apply(lb3d_ctx *c, int32_t stage, int32_t nt)
{
    g_err.clear();
    if (!c) return fail(LB3D_EINVAL, "NULL context");
    if (stage == 6) {
        if (c->state != IDLE) return fail(LB3D_ESTATE, "bc_apply(6) must follow a collision or init");
        if (c->le) {
            ensure_edges(c);
            launch_le_neighbours(c, nt);
            c->L->le_faces(c->b, false, compact_edges(c), c->stream);
            HIP_TRY(hipGetLastError());
        }
        c->nt = nt;
        c->state = AFTER_BC6;
    } else if (stage == 8) {
        if (c->state != AFTER_ADV) return fail(LB3D_ESTATE, "bc_apply(8) must follow advection");
        if (c->le) {
            c->L->le_faces(c->b, true, compact_edges(c), c->stream);
            HIP_TRY(hipGetLastError());
        }
        c->state = AFTER_BC8;
    } else if (stage == 10) {
        if (c->state != IDLE) return fail(LB3D_ESTATE, "bc_apply(10) must follow a collision");
    } else {
        return fail(LB3D_EINVAL, "bc_apply stage must be 6, 8 or 10");
    }
    return LB3D_OK;
}

int lb3d_gpu_advection(lb3d_ctx *c)
{
    g_err.clear();
    if (!c) return fail(LB3D_EINVAL, "NULL context");
    if (c->state != AFTER_BC6) return fail(LB3D_ESTATE, "advection must follow bc_apply(6)");
    // the one-pass plan streams inside the collision kernel
    if (!c->plan.fused) {
        c->L->moments(c->b, c->stream);
        HIP_TRY(hipGetLastError());
    }
    c->state = AFTER_ADV;
    return LB3D_OK;
}

// The collision of the current plan: f[cur] -> f[1-cur]; the ternary dipole
// relaxation reads the state's field db and writes db_alt (both plans).
// Plans are validated when they are set (lb3d_gpu_set_plan, autotune,
// create); a launcher that still rejects one returns LB3D_EINVAL.
static int launch_collision(lb3d_ctx *c)
{
    if (c->plan.fused) {
        if (!c->L->fused(c->b, c->plan.tx, c->plan.ty, c->plan.lz, c->plan.stage, c->stream))
            return fail(LB3D_EINVAL, "the one-pass plan has no kernel for this lattice");
    } else {
        c->L->dipole(c->b, c->stream);
        c->L->collide(c->b, c->stream);
    }
    return LB3D_OK;
}

static void mark_stepped(lb3d_ctx *c)
{
    c->aux_valid = !c->plan.fused;
    c->aux_from_prev = c->plan.fused;
}

int lb3d_gpu_collision(lb3d_ctx *c)
{
    g_err.clear();
    if (!c) return fail(LB3D_EINVAL, "NULL context");
    if (c->state != AFTER_BC8) return fail(LB3D_ESTATE, "collision must follow bc_apply(8)");
    int rc = launch_collision(c);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    flip(c);
    c->edges_ok = compact_edges(c);
    c->state = IDLE;
    mark_stepped(c);
    return LB3D_OK;
}

// One whole timestep on the stream (graph-capturable: kernels only).
static int enqueue_step(lb3d_ctx *c)
{
    // Lees-Edwards, two-pass plan: the halo-plane kernels compute the step's
    // offset state from the stored one and the collide stores it (no launch
    // of k_le_neighbours); the edge columns are valid (ensure_edges before
    // the steps are enqueued).  One-pass plan: k_le_neighbours, then the
    // halo planes from the lattice.
    const bool merged = c->le && !c->plan.fused;
    if (c->le) {
        if (!merged) launch_le_neighbours(c, -1);
        c->L->le_faces(c->b, false, compact_edges(c), c->stream, merged ? -1 : LE_STATE_AS_IS);
    }
    if (c->plan.fused) {
        if (c->le) c->L->le_faces(c->b, true, false, c->stream);
    } else {
        c->L->moments(c->b, c->stream);
        if (c->le) c->L->le_faces(c->b, true, true, c->stream, -1);
    }
    if (merged) c->L->set_le_commit(-1);
    int rc = launch_collision(c);
    if (merged) c->L->set_le_commit(LE_STATE_AS_IS);
    if (rc) return rc;
    flip(c);
    return LB3D_OK;
}

static int ensure_graph(lb3d_ctx *c)
{
    if (c->graph) return LB3D_OK;
    // two steps starting from cur = 0 (A -> B -> A)
    int saved = c->b.cur;
    set_cur(c, 0);
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(c);
    if (rc == LB3D_OK) rc = enqueue_step(c);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    set_cur(c, saved);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(LB3D_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        c->graph = nullptr;
        return fail(LB3D_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    return LB3D_OK;
}

static int sync_le_counter(lb3d_ctx *c)
{
    if (!c->le) return LB3D_OK;
    // the device step counter follows the host's nt
    int rc = pull_le(c);
    if (rc) return rc;
    if (c->host_le.nt != c->nt) {
        c->host_le.nt = c->nt;
        if ((rc = push_le(c))) return rc;
    }
    return LB3D_OK;
}

int lb3d_gpu_run_steps(lb3d_ctx *c, int32_t nsteps)
{
    g_err.clear();
    if (!c) return fail(LB3D_EINVAL, "NULL context");
    if (nsteps < 0) return fail(LB3D_EINVAL, "nsteps must be >= 0");
    if (c->state != IDLE) return fail(LB3D_ESTATE, "run_steps needs a completed timestep");
    if (nsteps == 0) return LB3D_OK;
    int rc = sync_le_counter(c);
    if (rc) return rc;
    if ((rc = ensure_graph(c))) return rc;
    ensure_edges(c);
    int left = nsteps;
    if (c->b.cur == 1) {
        if ((rc = enqueue_step(c))) return rc;
        left--;
    }
    for (; left >= 2; left -= 2) HIP_TRY(hipGraphLaunch(c->graph, c->stream));
    if (left == 1 && (rc = enqueue_step(c))) return rc;
    HIP_TRY(hipGetLastError());
    c->edges_ok = compact_edges(c);
    c->nt += nsteps;
    mark_stepped(c);
    return LB3D_OK;
}

int lb3d_gpu_run_steps_timed(lb3d_ctx *c, int32_t nsteps, double *ms)
{
    g_err.clear();
    if (!c || !ms) return fail(LB3D_EINVAL, "NULL argument");
    if (c->state != IDLE) return fail(LB3D_ESTATE, "run_steps needs a completed timestep");
    int rc = sync_le_counter(c);
    if (rc) return rc;
    if ((rc = ensure_graph(c))) return rc;
    Events ev(2);
    HIP_TRY(ev.create());
    hipEvent_t e0 = ev[0], e1 = ev[1];
    HIP_TRY(hipEventRecord(e0, c->stream));
    rc = lb3d_gpu_run_steps(c, nsteps);
    if (rc == LB3D_OK) {
        HIP_TRY(hipEventRecord(e1, c->stream));
        HIP_TRY(hipEventSynchronize(e1));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, e0, e1));
        *ms = t;
    }
    return rc;
}

// Autotuner: times each candidate plan on the context's own lattice.  Trial
// steps read f[cur] and write f[1-cur] without swapping, so the state is
// unchanged (the LE scalars are not advanced; the ternary dipole field is
// saved and restored around the trials).
int lb3d_gpu_autotune(lb3d_ctx *c, int32_t trial_steps, char *report, int32_t report_len)
{
    g_err.clear();
    if (!c) return fail(LB3D_EINVAL, "NULL context");
    if (c->state != IDLE) return fail(LB3D_ESTATE, "autotune needs a completed timestep");
    if (trial_steps < 1) trial_steps = 1;
    std::vector<Plan> cands;
    Plan two;
    two.fused = false;
    cands.push_back(two);
    if (c->p.ncomp == 3) {
        const int lzs[] = {8, 16, 32, 64};
        for (int k = 0; k < N_FUSED_SHAPES; k++)
            for (int lz : lzs) {
                if (!c->L->fused_supported(FUSED_SHAPES[k][0], FUSED_SHAPES[k][1], 2)) continue;
                if (lz > c->p.nz && lz != 8) continue;
                Plan q;
                q.fused = true;
                q.tx = FUSED_SHAPES[k][0];
                q.ty = FUSED_SHAPES[k][1];
                q.lz = lz < c->p.nz ? lz : c->p.nz;
                q.stage = 2;
                q.xg = 0;
                cands.push_back(q);
            }
    }
    if (c->p.ncomp == 2) {
        const int lzs[] = {16, 32, 64, 128, 256};
        for (int k = 0; k < N_FUSED_SHAPES; k++)
            for (int st = 0; st < 3; st++)
                for (int lz : lzs) {
                    if (!c->L->fused_supported(FUSED_SHAPES[k][0], FUSED_SHAPES[k][1], st)) continue;
                    if (lz > c->p.nz && lz != 16) continue;
                    Plan q;
                    q.fused = true;
                    q.tx = FUSED_SHAPES[k][0];
                    q.ty = FUSED_SHAPES[k][1];
                    q.lz = lz < c->p.nz ? lz : c->p.nz;
                    q.stage = st;
                    q.xg = 0;
                    cands.push_back(q);
                }
    }
    // the trials write f[1-cur] and the ternary dipole field db_alt, never
    // the state (f[cur], db); the guard puts back the plan if a trial fails
    struct Guard {
        lb3d_ctx *c;
        Plan saved;
        bool restore = true;
        ~Guard()
        {
            if (restore) {
                c->plan = saved;
                c->L->xg = saved.xg;
            }
        }
    } g{c, c->plan};
    // f[1-cur] and db_alt no longer hold the last step's input, so the
    // moments of the last advection cannot be rebuilt from them afterwards;
    // the edge columns are gathered for the trials and are stale after them
    c->aux_valid = false;
    c->aux_from_prev = false;
    if (c->le) {
        c->L->le_gather_edges(c->b, c->stream);
        c->edges_ok = false;
    }
    Events ev(2);
    HIP_TRY(ev.create());
    hipEvent_t e0 = ev[0], e1 = ev[1];
    double best = 1e300;
    Plan bestp = g.saved;
    std::string rep;
    auto time_plan = [&](const Plan &q, double &per) -> int {
        c->plan = q;
        c->L->xg = q.xg;
        auto trial = [&]() {
            // (the edge columns are those of the state, gathered before the
            // trials; the two-pass trials' collide overwrites them with
            // values of the trial, which are not the state's)
            if (c->le) c->L->le_faces(c->b, false, !q.fused, c->stream);
            if (!q.fused) c->L->moments(c->b, c->stream);
            if (c->le) c->L->le_faces(c->b, true, !q.fused, c->stream);
            return launch_collision(c);
        };
        if (int rc = trial()) return rc;  // warm-up
        HIP_TRY(hipEventRecord(e0, c->stream));
        for (int t = 0; t < trial_steps; t++)
            if (int rc = trial()) return rc;
        HIP_TRY(hipEventRecord(e1, c->stream));
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        per = ms / trial_steps;
        char buf[112];
        std::snprintf(buf, sizeof buf, "%s %dx%d lz=%d %s xg=%u %.4f ms\n", q.fused ? "fused" : "two-pass",
                      q.fused ? q.tx : 0, q.fused ? q.ty : 0, q.fused ? q.lz : 0,
                      q.fused ? STAGE_NAMES[q.stage] : "-", q.xg, per);
        rep += buf;
        if (per < best) {
            best = per;
            bestp = q;
        }
        return LB3D_OK;
    };
    double per = 0.0;
    for (const Plan &q : cands)
        if (int rc = time_plan(q, per)) return rc;
    // then the XCD chunking of the best plan
    const Plan top = bestp;
    for (uint32_t xg : {1u, 2u, 4u, 8u, 16u, 32u, 64u, 0u}) {
        if (xg == top.xg) continue;
        Plan q = top;
        q.xg = xg;
        if (int rc = time_plan(q, per)) return rc;
    }
    g.restore = false;
    c->plan = bestp;
    c->L->xg = bestp.xg;
    c->tuned = true;
    drop_graph(c);
    char buf[112];
    std::snprintf(buf, sizeof buf, "chosen: %s %dx%d lz=%d %s xg=%u %.4f ms\n", bestp.fused ? "fused" : "two-pass",
                  bestp.fused ? bestp.tx : 0, bestp.fused ? bestp.ty : 0, bestp.fused ? bestp.lz : 0,
                  bestp.fused ? STAGE_NAMES[bestp.stage] : "-", bestp.xg, best);
    rep += buf;
    if (report && report_len > 0) {
        std::strncpy(report, rep.c_str(), (size_t)report_len - 1);
        report[report_len - 1] = 0;
    }
    return LB3D_OK;
}

int lb3d_gpu_set_plan(lb3d_ctx *c, int32_t fused, int32_t tx, int32_t ty, int32_t lz, int32_t xcd_group)
{
    g_err.clear();
    if (!c) return fail(LB3D_EINVAL, "NULL context");
    if (xcd_group < 0) return fail(LB3D_EINVAL, "xcd_group must be >= 0");
    Plan q;
    q.xg = (uint32_t)xcd_group;
    q.fused = fused != 0;
    // fused = 1: LDS stage, 2: registers, 3: pulled again by the collide
    if (fused < 0 || fused > 3) return fail(LB3D_EINVAL, "fused must be 0..3");
    q.stage = fused == 1 ? 1 : (fused == 2 ? 0 : 2);
    if (q.fused) {
        if (c->p.ncomp == 2 && c->p.bdist != 2)
            return fail(LB3D_EINVAL, "the binary one-pass kernel needs bdist 2");
        if (c->p.ncomp == 3 && (fused != 3 || c->le))
            return fail(LB3D_EINVAL, "the ternary one-pass kernel is fused = 3 on periodic lattices");
        bool ok = c->L->fused_supported(tx, ty, q.stage);
        if (!ok || lz < 1) return fail(LB3D_EINVAL, "no such tile shape");
        q.tx = tx;
        q.ty = ty;
        q.lz = lz < c->p.nz ? lz : c->p.nz;
    }
    c->plan = q;
    c->L->xg = q.xg;
    c->edges_ok = false;
    drop_graph(c);
    return LB3D_OK;
}

int lb3d_gpu_synchronize(lb3d_ctx *c)
{
    g_err.clear();
    if (!c) return fail(LB3D_EINVAL, "NULL context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LB3D_OK;
}

int lb3d_gpu_get_density(lb3d_ctx *c, double *out)
{
    g_err.clear();
    if (!c || !out) return fail(LB3D_EINVAL, "NULL argument");
    const int C = c->p.ncomp;
    const size_t ns = (size_t)c->p.nx * c->p.ny * c->p.nz;
    // one component plane at a time keeps the host buffer small
    std::vector<double> plane(ns);
    std::vector<float> planef(is_dev(c) ? ns : 0);
    const size_t rs = c->L->real_bytes();
    for (int k = 0; k < C; k++) {
        double *rho = out + (size_t)k * ns;
        for (size_t s = 0; s < ns; s++) rho[s] = 0.0;
        for (int i = 0; i < 19; i++) {
            const char *src = (const char *)c->b.f[c->b.cur] + ((size_t)k * 19 + i) * ns * rs;
            if (is_dev(c)) {
                HIP_TRY(hipMemcpyAsync(planef.data(), src, ns * 4, hipMemcpyDeviceToHost, c->stream));
                HIP_TRY(hipStreamSynchronize(c->stream));
                double base = w_of(i) * rho0_of(c, k);
                for (size_t s = 0; s < ns; s++) rho[s] = rho[s] + (base + (double)planef[s]) * (double)host::GD[i];
            } else {
                HIP_TRY(hipMemcpyAsync(plane.data(), src, ns * 8, hipMemcpyDeviceToHost, c->stream));
                HIP_TRY(hipStreamSynchronize(c->stream));
                for (size_t s = 0; s < ns; s++) rho[s] = rho[s] + plane[s] * (double)host::GD[i];
            }
        }
    }
    return LB3D_OK;
}

int lb3d_gpu_get_momentum(lb3d_ctx *c, double *out)
{
    g_err.clear();
    if (!c || !out) return fail(LB3D_EINVAL, "NULL argument");
    const int C = c->p.ncomp;
    const size_t ns = (size_t)c->p.nx * c->p.ny * c->p.nz;
    std::vector<double> plane(ns);
    std::vector<float> planef(is_dev(c) ? ns : 0);
    const size_t rs = c->L->real_bytes();
    for (int k = 0; k < C; k++) {
        double *P = out + (size_t)k * 3 * ns;
        for (size_t s = 0; s < 3 * ns; s++) P[s] = 0.0;
        for (int i = 0; i < 19; i++) {
            const char *src = (const char *)c->b.f[c->b.cur] + ((size_t)k * 19 + i) * ns * rs;
            if (is_dev(c)) {
                HIP_TRY(hipMemcpyAsync(planef.data(), src, ns * 4, hipMemcpyDeviceToHost, c->stream));
                HIP_TRY(hipStreamSynchronize(c->stream));
                double base = w_of(i) * rho0_of(c, k);
                for (size_t s = 0; s < ns; s++) plane[s] = base + (double)planef[s];
            } else {
                HIP_TRY(hipMemcpyAsync(plane.data(), src, ns * 8, hipMemcpyDeviceToHost, c->stream));
                HIP_TRY(hipStreamSynchronize(c->stream));
            }
            // P += g_i n_i c_i, i ascending (lb3d_update_velocity)
            const double g = (double)host::GD[i];
            const int cv[3] = {host::CX[i], host::CY[i], host::CZ[i]};
            for (int a = 0; a < 3; a++) {
                if (cv[a] == 0) continue;
                double *Pa = P + (size_t)a * ns;
                for (size_t s = 0; s < ns; s++) Pa[s] = Pa[s] + (plane[s] * g) * (double)cv[a];
            }
        }
    }
    return LB3D_OK;
}

// The reference's arr observable (dump_arrows, code/lb3d_io_dump_data.F90:
// 1308-1460): per site the momentum of every component less half its
// Shan-Chen force times its density, weighted by amass, over the
// amass-weighted density (max with 10.e-9 in single precision, as the
// reference's real() and single literal give), less half the body-force
// acceleration inside its box.  The forces come from the device (the
// collide's own force code, from the moments of the last advection and the
// dipole field); the sums are FP64 in the reference's order.
int lb3d_gpu_get_arrows(lb3d_ctx *c, double *out)
{
    g_err.clear();
    if (!c || !out) return fail(LB3D_EINVAL, "NULL argument");
    if (c->state != IDLE) return fail(LB3D_ESTATE, "get_arrows needs a completed timestep");
    int rc = rebuild_aux(c);
    if (rc) return rc;
    if (!c->aux_valid)
        return fail(LB3D_ESTATE, "get_arrows needs the moments of an advection: run a timestep first");
    const int C = c->p.ncomp, nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
    const size_t ns = (size_t)nx * ny * nz, rs = c->L->real_bytes();
    std::vector<double> pops((size_t)C * 19 * ns), F((size_t)C * 3 * ns);
    if ((rc = lb3d_gpu_get_populations(c, pops.data()))) return rc;
    {
        // a temporary device buffer of the force field (not a hot path)
        void *d = nullptr;
        HIP_TRY(hipMalloc(&d, F.size() * rs));
        c->L->forces(c->b, d, c->stream);
        std::vector<unsigned char> h(F.size() * rs);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(h.data(), d, h.size(), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        (void)hipFree(d);
        if (e != hipSuccess) return fail(LB3D_EHIP, std::string("forces: ") + hipGetErrorString(e));
        for (size_t q = 0; q < F.size(); q++)
            F[q] = rs == 8 ? ((const double *)h.data())[q] : (double)((const float *)h.data())[q];
    }
    std::vector<uint8_t> rock;
    if (c->b.rock) {
        rock.resize(ns);
        HIP_TRY(hipMemcpyAsync(rock.data(), c->b.rock, ns, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    const lb3d_params &p = c->p;
    const bool body = p.g_accn[0] != 0.0 || p.g_accn[1] != 0.0 || p.g_accn[2] != 0.0;
    const int ext[3] = {nx, ny, nz};
    int bmax[3];
    for (int k = 0; k < 3; k++) bmax[k] = p.g_accn_max[k] == 0 ? ext[k] : p.g_accn_max[k];
    const int *cv[3] = {host::CX, host::CY, host::CZ};
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const size_t s = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z);
                const bool fluid = rock.empty() || !rock[s];
                double rho[3] = {0, 0, 0}, mom[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                for (int k = 0; k < C; k++)
                    for (int i = 0; i < 19; i++) {
                        const double v = pops[((size_t)k * 19 + i) * ns + s] * (double)host::GD[i];
                        rho[k] = rho[k] + v;
                        for (int a = 0; a < 3; a++) mom[k][a] = mom[k][a] + v * (double)cv[a][i];
                    }
                double den = rho[0] * p.amass[0];
                for (int k = 1; k < C; k++) den = den + rho[k] * p.amass[k];
                const float denf = (float)den;
                const double dd = (double)(denf > 10.e-9f ? denf : 10.e-9f);
                const bool inbox = body && fluid && z + 1 >= p.g_accn_min[2] && z + 1 <= bmax[2] &&
                                   x + 1 >= p.g_accn_min[0] && x + 1 <= bmax[0] && y + 1 >= p.g_accn_min[1] &&
                                   y + 1 <= bmax[1];
                for (int a = 0; a < 3; a++) {
                    double v = 0.0;
                    for (int k = 0; k < C; k++) {
                        const double fa = fluid ? F[((size_t)k * 3 + a) * ns + s] : 0.0;
                        const double A = mom[k][a] * p.amass[k];
                        const double B = ((fa / 2.0) * rho[k]) * p.amass[k];
                        v = k == 0 ? A - B : (v + A) - B;
                    }
                    v = v / dd;
                    if (inbox) v = v - p.g_accn[a] / 2.0;
                    out[3 * s + a] = v;
                }
            }
    return LB3D_OK;
}

int lb3d_gpu_le_halo(lb3d_ctx *c, double *out)
{
    g_err.clear();
    if (!c || !out) return fail(LB3D_EINVAL, "NULL argument");
    if (!c->le) return fail(LB3D_EINVAL, "not a Lees-Edwards lattice");
    if (c->state != IDLE) return fail(LB3D_ESTATE, "le_halo needs a completed timestep");
    int rc = sync_le_counter(c);
    if (rc) return rc;
    // the stage-6 x-halo planes of the current populations and offset
    ensure_edges(c);
    c->L->le_faces(c->b, false, compact_edges(c), c->stream);
    HIP_TRY(hipGetLastError());
    const size_t plane = (size_t)c->p.ny * c->p.nz;
    const size_t n = 2 * 2 * 19 * plane;
    const size_t rs = c->L->real_bytes();
    if (is_dev(c)) {
        std::vector<float> h(n);
        HIP_TRY(hipMemcpyAsync(h.data(), c->b.face_pc, n * rs, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (size_t f = 0; f < 2; f++)
            for (int k = 0; k < 2; k++)
                for (int i = 0; i < 19; i++) {
                    size_t o = ((f * 2 + k) * 19 + i) * plane;
                    double base = w_of(i) * rho0_of(c, k);
                    for (size_t t = 0; t < plane; t++) out[o + t] = base + (double)h[o + t];
                }
    } else {
        HIP_TRY(hipMemcpyAsync(out, c->b.face_pc, n * rs, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return LB3D_OK;
}

int lb3d_gpu_get_nt(lb3d_ctx *c, int32_t *nt)
{
    if (!c || !nt) return fail(LB3D_EINVAL, "NULL argument");
    *nt = c->nt;
    return LB3D_OK;
}

int lb3d_gpu_diagnostics(lb3d_ctx *c, double mass[3], double momentum[3], int64_t counts[3])
{
    g_err.clear();
    if (!c || !mass || !momentum || !counts) return fail(LB3D_EINVAL, "NULL argument");
    unsigned nb = c->L->diag_blocks();
    c->L->diagnostics(c->b, c->d_partials, c->stream);
    HIP_TRY(hipGetLastError());
    std::vector<double> h((size_t)nb * NDIAG);
    HIP_TRY(hipMemcpyAsync(h.data(), c->d_partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double acc[NDIAG] = {0};
    for (unsigned b = 0; b < nb; b++)
        for (int k = 0; k < NDIAG; k++) acc[k] = acc[k] + h[(size_t)b * NDIAG + k];
    for (int k = 0; k < 3; k++) {
        mass[k] = k < c->p.ncomp ? acc[k] : 0.0;
        momentum[k] = acc[3 + k];
        counts[k] = (int64_t)llround(acc[6 + k]);
    }
    return LB3D_OK;
}


int lb3d_gpu_sanity(lb3d_ctx *c, lb3d_sanity *out)
{
    g_err.clear();
    if (!c || !out) return fail(LB3D_EINVAL, "NULL argument");
    unsigned nb = c->L->diag_blocks();
    c->L->sanity(c->b, c->d_partials, c->stream);
    HIP_TRY(hipGetLastError());
    std::vector<double> h((size_t)nb * NSAN);
    HIP_TRY(hipMemcpyAsync(h.data(), c->d_partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    reduce_sanity(h.data(), nb, c->p.ncomp, c->p.nx, c->p.ny, c->p.nz, out);
    return LB3D_OK;
}

int lb3d_gpu_le_state(lb3d_ctx *c, double *shear_sum, int32_t *lelo, double *lefrac, double *shear_u)
{
    g_err.clear();
    if (!c || !shear_sum || !lelo || !lefrac || !shear_u) return fail(LB3D_EINVAL, "NULL argument");
    if (c->le) {
        int rc = pull_le(c);
        if (rc) return rc;
    }
    *shear_sum = c->host_le.shear_sum;
    *lelo = c->host_le.lelo;
    *lefrac = c->host_le.lefrac;
    *shear_u = c->host_le.shear_u;
    return LB3D_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// lb3d_gpu_step_noalloc: one timestep over caller-owned device buffers that
// allocates nothing.  The moments field of the two-pass step (the scratch
// of lb3d_gpu_run_steps' kernels) lives in a file-scope device array of
// LB3D_NOALLOC_SCRATCH_BYTES, which every GPU the library runs on holds
// until the process ends.  In host code `scratch` names a host-side stand-in;
// its device address is looked up once per device and used instead.
static __device__ __align__(256) unsigned char scratch[LB3D_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch[MAX_DEVICES] = {};

// The device address of `scratch` on the current device.
int scratch_on_device(void **out)
{
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(LB3D_EINVAL, "device index beyond 64");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(LB3D_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = ptr;
    }
    *out = g_scratch[dev];
    return LB3D_OK;
}
}  // namespace

extern "C" int lb3d_gpu_step_noalloc(const lb3d_params *p, const void *f_in, void *f_out, const void *db_in,
                                     void *db_out, void *stream)
{
    g_err.clear();
    int rc = check_params(p);
    if (rc) return rc;
    if (!f_in || !f_out || f_in == f_out) return fail(LB3D_EINVAL, "f_in and f_out must be distinct device buffers");
    if (p->boundary != LB3D_BC_PERIODIC)
        return fail(LB3D_EINVAL, "lb3d_gpu_step_noalloc runs periodic lattices (Lees-Edwards needs face planes)");
    if (p->ncomp == 3 && (!db_in || !db_out || db_in == db_out))
        return fail(LB3D_EINVAL, "ternary lattices need distinct db_in and db_out device buffers");
    const uint64_t ns = (uint64_t)p->nx * p->ny * p->nz;
    const size_t rs = p->precision == LB3D_FP64 ? 8 : 4;
    const uint64_t need = aux_elems(p->ncomp, ns) * rs;
    if (need > LB3D_NOALLOC_SCRATCH_BYTES) {
        char m[200];
        std::snprintf(m, sizeof m,
                      "lattice needs %llu bytes of scratch, lb3d_gpu_step_noalloc has %llu: use lb3d_gpu_create "
                      "and lb3d_gpu_run_steps",
                      (unsigned long long)need, (unsigned long long)LB3D_NOALLOC_SCRATCH_BYTES);
        return fail(LB3D_EINVAL, m);
    }
    void *aux = nullptr;
    if ((rc = scratch_on_device(&aux))) return rc;
    DevBufs b;
    b.f[0] = const_cast<void *>(f_in);
    b.f[1] = f_out;
    b.cur = 0;
    b.aux = aux;  // written in full by the moments pass before anything reads it
    b.db = const_cast<void *>(db_in);
    b.db_alt = db_out;
    hipStream_t s = (hipStream_t)stream;
    if (p->precision == LB3D_FP64)
        step_buffers_fp64(*p, b, s);
    else
        step_buffers_fp32(*p, b, s);
    HIP_TRY(hipGetLastError());
    return LB3D_OK;
}

extern "C" {

int lb3d_gpu_profile_steps(lb3d_ctx *c, int32_t nsteps, int32_t max_kernels, const char **names, double *ms,
                           int32_t *nkernels, double *ms_per_step)
{
    g_err.clear();
    if (!c || !names || !ms || !nkernels || !ms_per_step) return fail(LB3D_EINVAL, "NULL argument");
    if (nsteps < 1 || max_kernels < K_COUNT) return fail(LB3D_EINVAL, "nsteps >= 1 and room for 6 kernels needed");
    if (c->state != IDLE) return fail(LB3D_ESTATE, "profile_steps needs a completed timestep");
    int rc0 = sync_le_counter(c);
    if (rc0) return rc0;
    Events ev((size_t)nsteps * K_COUNT * 2 + 2);
    std::vector<int> used((size_t)nsteps * K_COUNT, 0);
    HIP_TRY(ev.create());
    hipEvent_t t0 = ev[ev.e.size() - 2], t1 = ev[ev.e.size() - 1];
    auto rec = [&](int step, int k, bool start) {
        return hipEventRecord(ev[((size_t)step * K_COUNT + k) * 2 + (start ? 0 : 1)], c->stream);
    };
    HIP_TRY(hipEventRecord(t0, c->stream));
    for (int st = 0; st < nsteps; st++) {
        auto mark = [&](int k) { used[(size_t)st * K_COUNT + k] = 1; };
        // as enqueue_step: the two-pass plan advances the Lees-Edwards state
        // in its halo-plane kernels and collide, the one-pass plan launches
        // k_le_neighbours
        const bool merged = c->le && !c->plan.fused;
        if (c->le) {
            if (!merged) {
                HIP_TRY(rec(st, K_LEN, true));
                launch_le_neighbours(c, -1);
                HIP_TRY(rec(st, K_LEN, false));
                mark(K_LEN);
            }
            HIP_TRY(rec(st, K_LE6, true));
            ensure_edges(c);
            c->L->le_faces(c->b, false, compact_edges(c), c->stream, merged ? -1 : LE_STATE_AS_IS);
            HIP_TRY(rec(st, K_LE6, false));
            mark(K_LE6);
        }
        if (!c->plan.fused) {
            HIP_TRY(rec(st, K_MOM, true));
            c->L->moments(c->b, c->stream);
            HIP_TRY(rec(st, K_MOM, false));
            mark(K_MOM);
        }
        if (c->le) {
            HIP_TRY(rec(st, K_LE8, true));
            c->L->le_faces(c->b, true, compact_edges(c), c->stream, merged ? -1 : LE_STATE_AS_IS);
            HIP_TRY(rec(st, K_LE8, false));
            mark(K_LE8);
        }
        if (c->plan.fused) {
            HIP_TRY(rec(st, K_FUS, true));
            if (!c->L->fused(c->b, c->plan.tx, c->plan.ty, c->plan.lz, c->plan.stage, c->stream))
                return fail(LB3D_EINVAL, "the one-pass plan has no kernel for this lattice");
            HIP_TRY(rec(st, K_FUS, false));
            mark(K_FUS);
        } else {
            if (c->p.ncomp == 3) {
                HIP_TRY(rec(st, K_DIP, true));
                c->L->dipole(c->b, c->stream);
                HIP_TRY(rec(st, K_DIP, false));
                mark(K_DIP);
            }
            HIP_TRY(rec(st, K_COL, true));
            if (merged) c->L->set_le_commit(-1);
            c->L->collide(c->b, c->stream);
            if (merged) c->L->set_le_commit(LE_STATE_AS_IS);
            HIP_TRY(rec(st, K_COL, false));
            mark(K_COL);
        }
        flip(c);
        c->edges_ok = compact_edges(c);
    }
    HIP_TRY(hipEventRecord(t1, c->stream));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    double tot[K_COUNT] = {0};
    int cnt[K_COUNT] = {0};
    for (int st = 0; st < nsteps; st++)
        for (int k = 0; k < K_COUNT; k++)
            if (used[(size_t)st * K_COUNT + k]) {
                float m = 0.f;
                HIP_TRY(hipEventElapsedTime(&m, ev[((size_t)st * K_COUNT + k) * 2], ev[((size_t)st * K_COUNT + k) * 2 + 1]));
                tot[k] += m;
                cnt[k]++;
            }
    float all = 0.f;
    HIP_TRY(hipEventElapsedTime(&all, t0, t1));
    int nk = 0;
    for (int k = 0; k < K_COUNT; k++)
        if (cnt[k]) {
            names[nk] = KNAMES[k];
            ms[nk] = tot[k] / cnt[k];
            nk++;
        }
    *nkernels = nk;
    *ms_per_step = all / nsteps;
    c->nt += nsteps;
    mark_stepped(c);
    return LB3D_OK;
}

}  // extern "C"
