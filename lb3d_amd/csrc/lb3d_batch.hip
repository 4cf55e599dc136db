// lb3d_batch.hip -- the batched ensemble of include/lb3d_gpu.h: many
// independent periodic lattices of uneven sizes (parameter sweeps, seeds)
// advanced together.  All lattices live in slices of shared buffers; one
// timestep is the moments, dipole and collide phases, each a single
// persistent launch over a work list of (lattice, 256-site tile) items, and
// whole steps are replayed from a captured HIP graph.  The persistent grid
// size is chosen by lb3d_gpu_batch_autotune.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "lb3d_kernels.h"
#include "lb3d_solver.h"

using namespace lb3d;

#define BTRY(expr)                                                                          \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return set_error(LB3D_EHIP, (std::string(#expr) + ": " + hipGetErrorString(e_)).c_str()); \
    } while (0)

struct lb3d_batch {
    int nlat = 0, C = 2, prec = LB3D_FP64;
    std::vector<lb3d_params> p;
    std::vector<Launcher *> L;  // per-lattice kernels (init, diagnostics, sequential plan)
    BatchLauncher *B = nullptr;
    size_t rs = 8;
    std::vector<uint64_t> f_off, aux_off, db_off, ns;
    void *f[2] = {nullptr, nullptr};
    void *aux = nullptr, *db = nullptr, *lat = nullptr, *items = nullptr;
    double *partials = nullptr;
    uint32_t nitems = 0;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int cur = 0, nt = 0;
    unsigned nwg = 2048;     // persistent grid size, a multiple of 8
    uint32_t xg = 0;         // work items per XCD chunk (xcd_tile)
    bool sequential = false; // one lattice after another with the single-lattice kernels
    // grouped persistent plan: consecutive lattices whose populations add up
    // to about group_bytes run their three phases before the next group, so
    // that the collide re-reads populations and moments that the moments
    // phase has just brought into the memory-side cache (0: one group)
    uint64_t group_bytes = 0;
    std::vector<uint32_t> item0;   // first work item of each lattice, then nitems
    std::vector<uint32_t> groups;  // first work item of each group, then nitems
    // collide-moments plan (FP32 ternary, lb3d_gpu_batch_set_collide_moments):
    // tiles of about cm_cols columns and z-chunks of about cm_lz planes
    int cm_cols = 0;  // 0: the plan is off
    int cm_lz = 0;
    int cm_threads = 256;  // work-group size of k_batch_cm (256 or 512)
    void *aux1 = nullptr;  // second aux field: aux(j) holds the moments of f[j]'s streamed state
    void *cm_items = nullptr, *edge_items = nullptr, *edge_sites = nullptr;
    uint32_t cm_nitems = 0, edge_nitems = 0;
    uint64_t cm_edge_sites = 0;  // edge sites (moments pulled by k_batch_edge_moments)
    bool aux_ok = false;         // aux(cur) holds the moments of the state's streamed populations
};

namespace {

uint64_t round64(uint64_t v) { return (v + 63) & ~uint64_t(63); }

DevBufs view(const lb3d_batch *b, int k)
{
    DevBufs v;
    for (int j = 0; j < 2; j++) v.f[j] = (char *)b->f[j] + b->f_off[k] * b->rs;
    v.aux = (char *)b->aux + b->aux_off[k] * b->rs;
    // one dipole field per lattice, relaxed in place: the dipole pass writes
    // db_alt after the moments pass has read db, and the collide reads db_alt
    v.db = b->C == 3 ? (char *)b->db + b->db_off[k] * b->rs : nullptr;
    v.db_alt = v.db;
    v.cur = b->cur;
    return v;
}

// Split the lattices into groups of consecutive lattices holding about
// group_bytes of populations each (a larger lattice forms its own group).
void make_groups(lb3d_batch *b)
{
    b->groups.assign(1, 0u);
    if (b->group_bytes) {
        uint64_t acc = 0;
        for (int k = 0; k < b->nlat; k++) {
            uint64_t by = (uint64_t)b->C * 19u * b->ns[k] * b->rs;
            if (acc > 0 && acc + by > b->group_bytes) {
                b->groups.push_back(b->item0[k]);
                acc = 0;
            }
            acc += by;
        }
    }
    b->groups.push_back(b->nitems);
}

// Phase ph of the persistent plan over the work items of group g.
void group_phase(lb3d_batch *b, size_t g, int ph)
{
    const uint32_t i0 = b->groups[g], n = b->groups[g + 1] - i0;
    const unsigned nwg = b->nwg < xcd_round(n) ? b->nwg : xcd_round(n);
    b->B->phase(ph, b->lat, (const uint2 *)b->items + i0, n, nwg, b->f[b->cur], b->f[1 - b->cur], b->aux, b->db,
                b->C, b->stream, true);
}

void *aux_of(const lb3d_batch *b, int j) { return j == 0 ? b->aux : b->aux1; }

// The moments of the state's streamed populations into aux(cur), which the
// collide-moments plan carries from step to step.
void prime_aux(lb3d_batch *b)
{
    b->B->phase(0, b->lat, b->items, b->nitems, xcd_round(b->nitems), b->f[b->cur], b->f[1 - b->cur],
                aux_of(b, b->cur), b->db, b->C, b->stream, false);
    b->aux_ok = true;
}

// One step of the collide-moments plan: dipole relaxation from aux(cur),
// collision with the gathered moments into aux(1-cur), then the moments of
// the tiles' edge sites.  Phases ph < 0: all three; else only that one
// (0 edge moments, 1 dipole, 2 collide) without swapping.
void cm_phase(lb3d_batch *b, int ph)
{
    const int c = b->cur;
    if (ph == 1)
        b->B->phase(1, b->lat, b->items, b->nitems, xcd_round(b->nitems), b->f[c], b->f[1 - c], aux_of(b, c), b->db,
                    b->C, b->stream, false);
    else if (ph == 2)
        b->B->cm_collide(b->lat, b->cm_items, b->cm_nitems, xcd_round(b->cm_nitems), b->cm_threads, b->f[c],
                         b->f[1 - c], aux_of(b, c), aux_of(b, 1 - c), b->db, b->stream);
    else
        b->B->edge_moments(b->lat, b->edge_items, b->edge_nitems, xcd_round(b->edge_nitems), b->edge_sites,
                           b->f[1 - c], b->db, aux_of(b, 1 - c), b->stream);
}

void enqueue_batch_step(lb3d_batch *b)
{
    if (b->cm_cols) {
        cm_phase(b, 1);
        cm_phase(b, 2);
        cm_phase(b, 0);
        b->cur = 1 - b->cur;
        return;
    }
    b->aux_ok = false;  // the other plans use aux(0) as scratch
    if (b->sequential) {
        for (int k = 0; k < b->nlat; k++) {
            DevBufs v = view(b, k);
            b->L[k]->moments(v, b->stream);
            b->L[k]->dipole(v, b->stream);
            b->L[k]->collide(v, b->stream);
        }
    } else {
        for (size_t g = 0; g + 1 < b->groups.size(); g++)
            for (int ph = 0; ph < 3; ph++) group_phase(b, g, ph);
    }
    b->cur = 1 - b->cur;
}

int drop_graph(lb3d_batch *b)
{
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    return LB3D_OK;
}

// Work items and edge-site lists of the collide-moments plan.  Per
// lattice: kx tiles of tx columns (tx + 2 <= threads, the region's row
// stride), rows ty = threads / (tx + 2) evened out over the lattice,
// z-chunks of about lz planes.  Lattices go largest first, each z-chunk's
// tiles consecutively (the tiles that share halos run together on one XCD).
int build_cm(lb3d_batch *b, int cols, int lz, int threads)
{
    if (threads != 256 && threads != 512) return set_error(LB3D_EINVAL, "tile_threads must be 256 or 512");
    if (!b->B->cm_supported(b->C))
        return set_error(LB3D_EINVAL, "the collide-moments plan runs FP32 ternary batches of bdist 0 and 2");
    if (cols < 1 || cols > threads - 2 || lz < 1)
        return set_error(LB3D_EINVAL, "need 1 <= tile_cols <= tile_threads - 2 and lz >= 1");
    b->cm_cols = 0;
    b->cm_edge_sites = 0;
    std::vector<int> order(b->nlat);
    for (int k = 0; k < b->nlat; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b->ns[a] > b->ns[c]; });
    std::vector<uint4> items;
    std::vector<uint32_t> sites;
    std::vector<uint2> eitems;
    for (int k : order) {
        const lb3d_params &p = b->p[k];
        const int nx = p.nx, ny = p.ny, nz = p.nz;
        if (nx > 65535 || ny > 65535 || nz > 65535) return set_error(LB3D_EINVAL, "lattice edge above 65535");
        int kx = (nx + cols / 2) / cols;
        if (kx < 1) kx = 1;
        int tx = (nx + kx - 1) / kx;
        while (tx + 2 > threads) tx = (nx + ++kx - 1) / kx;
        kx = (nx + tx - 1) / tx;
        const int S = tx + 2;
        int ty = threads / S < ny ? threads / S : ny;
        int ky = (ny + ty - 1) / ty;
        ty = (ny + ky - 1) / ky;
        ky = (ny + ty - 1) / ty;
        int kz = (nz + lz / 2) / lz;
        if (kz < 1) kz = 1;
        int lzc = (nz + kz - 1) / kz;
        kz = (nz + lzc - 1) / lzc;
        std::vector<uint32_t> es;
        for (int zc = 0; zc < kz; zc++)
            for (int yt = 0; yt < ky; yt++)
                for (int xt = 0; xt < kx; xt++) {
                    const int x0 = xt * tx, y0 = yt * ty, z0 = zc * lzc;
                    const int tw = nx - x0 < tx ? nx - x0 : tx, th = ny - y0 < ty ? ny - y0 : ty;
                    const int z1 = nz - z0 < lzc ? nz : z0 + lzc;
                    items.push_back(make_uint4((unsigned)k, (unsigned)x0 | (unsigned)y0 << 16,
                                               (unsigned)z0 | (unsigned)z1 << 16,
                                               (unsigned)tw | (unsigned)th << 9 | (unsigned)S << 18));
                    for (int z = z0; z < z1; z++)
                        for (int r = 0; r < th; r++) {
                            const bool edge_row = r == 0 || r == th - 1 || z == z0 || z == z1 - 1;
                            if (!edge_row) continue;
                            for (int c = 0; c < tw; c++)
                                es.push_back((uint32_t)(x0 + c) + (uint32_t)nx * ((uint32_t)(y0 + r) + (uint32_t)ny * z));
                        }
                }
        std::sort(es.begin(), es.end());
        b->cm_edge_sites += es.size();
        for (size_t q = 0; q < es.size(); q += BLOCK) eitems.push_back(make_uint2((unsigned)k, (unsigned)(sites.size() + q)));
        es.resize((es.size() + BLOCK - 1) / BLOCK * BLOCK, ~0u);
        sites.insert(sites.end(), es.begin(), es.end());
    }
    for (void **q : {&b->cm_items, &b->edge_items, &b->edge_sites})
        if (*q) {
            (void)hipFree(*q);
            *q = nullptr;
        }
    b->cm_nitems = b->edge_nitems = 0;
    auto up = [&](void **dst, const void *src, size_t bytes) -> int {
        BTRY(hipMalloc(dst, bytes ? bytes : 16));
        if (bytes) BTRY(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return LB3D_OK;
    };
    if (int rc = up(&b->cm_items, items.data(), items.size() * sizeof(uint4))) return rc;
    if (int rc = up(&b->edge_items, eitems.data(), eitems.size() * sizeof(uint2))) return rc;
    if (int rc = up(&b->edge_sites, sites.data(), sites.size() * sizeof(uint32_t))) return rc;
    if (!b->aux1) {
        const uint64_t elems = b->aux_off.back() + aux_elems(b->C, b->ns.back());
        BTRY(hipMalloc(&b->aux1, elems * b->rs));
    }
    b->cm_nitems = (uint32_t)items.size();
    b->edge_nitems = (uint32_t)eitems.size();
    b->cm_cols = cols;
    b->cm_lz = lz;
    b->cm_threads = threads;
    b->aux_ok = false;
    return LB3D_OK;
}

int ensure_graph(lb3d_batch *b)
{
    if (b->graph) return LB3D_OK;
    int saved = b->cur;
    b->cur = 0;
    hipGraph_t g = nullptr;
    BTRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    enqueue_batch_step(b);
    enqueue_batch_step(b);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    b->cur = saved;
    if (e != hipSuccess) return set_error(LB3D_EHIP, (std::string("hipStreamEndCapture: ") + hipGetErrorString(e)).c_str());
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->graph = nullptr;
        return set_error(LB3D_EHIP, (std::string("hipGraphInstantiate: ") + hipGetErrorString(e)).c_str());
    }
    return LB3D_OK;
}

}  // namespace

extern "C" {

int lb3d_gpu_batch_destroy(lb3d_batch *b);

int lb3d_gpu_batch_create(int32_t nlat, const lb3d_params *params, lb3d_batch **out)
{
    clear_error();
    if (!out || !params || nlat < 1) return set_error(LB3D_EINVAL, "need nlat >= 1, params and out");
    *out = nullptr;
    const lb3d_params &p0 = params[0];
    for (int k = 0; k < nlat; k++) {
        const lb3d_params &q = params[k];
        if (q.ncomp != p0.ncomp || q.precision != p0.precision)
            return set_error(LB3D_EINVAL, "all lattices of a batch share ncomp and precision");
        if (q.boundary != LB3D_BC_PERIODIC) return set_error(LB3D_EINVAL, "batched lattices are periodic");
        if (q.ncomp != 2 && q.ncomp != 3) return set_error(LB3D_EINVAL, "ncomp must be 2 or 3");
        if (q.bdist < 0 || q.bdist > 6) return set_error(LB3D_EINVAL, "bdist must be 0..6");
        if (q.inv_fluid != 0) return set_error(LB3D_EINVAL, "the batch runs lattices without invasion");
        if (q.precision != LB3D_FP64 && q.precision != LB3D_FP32_DEV)
            return set_error(LB3D_EINVAL, "precision must be LB3D_FP64 or LB3D_FP32_DEV");
        if (q.nx < 2 || q.ny < 1 || q.nz < 1) return set_error(LB3D_EINVAL, "lattice extent must be nx>=2, ny>=1, nz>=1");
        for (int c = 0; c < q.ncomp; c++)
            if (!(q.tau[c] > 0.0) || !(q.amass[c] > 0.0)) return set_error(LB3D_EINVAL, "tau and amass must be positive");
        if (q.ncomp == 3 && !(q.tau_d > 0.0)) return set_error(LB3D_EINVAL, "tau_d must be positive");
    }
    lb3d_batch *b = new lb3d_batch();
    b->nlat = nlat;
    b->C = p0.ncomp;
    b->prec = p0.precision;
    b->rs = b->prec == LB3D_FP64 ? 8 : 4;
    b->B = b->prec == LB3D_FP64 ? make_batch_launcher_fp64() : make_batch_launcher_fp32();
    uint64_t fo = 0, ao = 0, dbo = 0, maxns = 0;
    std::vector<uint2> items;
    for (int k = 0; k < nlat; k++) {
        b->p.push_back(params[k]);
        b->L.push_back(b->prec == LB3D_FP64 ? make_launcher_fp64(params[k]) : make_launcher_fp32(params[k]));
        uint64_t n = (uint64_t)params[k].nx * params[k].ny * params[k].nz;
        b->ns.push_back(n);
        b->f_off.push_back(fo);
        b->aux_off.push_back(ao);
        b->db_off.push_back(dbo);
        fo = round64(fo + (uint64_t)b->C * 19 * n);
        ao = round64(ao + aux_elems(b->C, n));
        dbo = round64(dbo + (b->C == 3 ? 3 * n : 0));
        if (n > maxns) maxns = n;
        if ((uint64_t)b->C * 19 * n >= 4294967296ull) {
            lb3d_gpu_batch_destroy(b);
            return set_error(LB3D_EINVAL, "a lattice is too large for 32-bit population indexing");
        }
        b->item0.push_back((uint32_t)items.size());
        for (uint64_t s0 = 0; s0 < n; s0 += BLOCK) items.push_back(make_uint2((unsigned)k, (unsigned)s0));
    }
    b->nitems = (uint32_t)items.size();
    b->item0.push_back(b->nitems);
    make_groups(b);
    int mask = 0;
    for (int k = 0; k < nlat; k++) mask |= 1 << params[k].bdist;
    b->B->set_bdist_mask(mask);
    b->B->set_pop_bytes(fo * b->rs);
    std::vector<unsigned char> lat((size_t)nlat * b->B->lat_bytes());
    for (int k = 0; k < nlat; k++)
        b->B->fill_lat(lat.data() + (size_t)k * b->B->lat_bytes(), params[k], b->f_off[k], b->aux_off[k], b->db_off[k]);
    auto fail_destroy = [&](const char *what, hipError_t e) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        lb3d_gpu_batch_destroy(b);
        return set_error(LB3D_ENOMEM, m.c_str());
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return fail_destroy("hipStreamCreate", e);
    for (int j = 0; j < 2; j++)
        if ((e = hipMalloc(&b->f[j], fo * b->rs)) != hipSuccess) return fail_destroy("hipMalloc populations", e);
    if ((e = hipMalloc(&b->aux, ao * b->rs)) != hipSuccess) return fail_destroy("hipMalloc aux", e);
    if (b->C == 3 && (e = hipMalloc(&b->db, dbo * b->rs)) != hipSuccess) return fail_destroy("hipMalloc db", e);
    if ((e = hipMalloc(&b->lat, lat.size())) != hipSuccess) return fail_destroy("hipMalloc table", e);
    if ((e = hipMalloc(&b->items, items.size() * sizeof(uint2))) != hipSuccess) return fail_destroy("hipMalloc items", e);
    if ((e = hipMalloc((void **)&b->partials, ((maxns + BLOCK - 1) / BLOCK) * NSAN * sizeof(double))) != hipSuccess)
        return fail_destroy("hipMalloc partials", e);
    if ((e = hipMemcpy(b->lat, lat.data(), lat.size(), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(b->items, items.data(), items.size() * sizeof(uint2), hipMemcpyHostToDevice)) != hipSuccess)
        return fail_destroy("hipMemcpy", e);
    b->nwg = xcd_round(b->nitems < 2048 ? b->nitems : 2048);
    *out = b;
    return LB3D_OK;
}

int lb3d_gpu_batch_destroy(lb3d_batch *b)
{
    if (!b) return LB3D_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    for (int j = 0; j < 2; j++)
        if (b->f[j]) (void)hipFree(b->f[j]);
    for (void *q : {b->aux, b->aux1, b->db, b->lat, b->items, (void *)b->partials, b->cm_items, b->edge_items,
                    b->edge_sites})
        if (q) (void)hipFree(q);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    for (Launcher *l : b->L) delete l;
    delete b->B;
    delete b;
    return LB3D_OK;
}

int lb3d_gpu_batch_init_noise(lb3d_batch *b, int32_t k, uint64_t seed, const double rho_mean[3],
                              const double amp[3], int32_t dipoles)
{
    clear_error();
    if (!b || k < 0 || k >= b->nlat || !rho_mean || !amp) return set_error(LB3D_EINVAL, "bad argument");
    b->L[k]->init_noise(view(b, k), seed, rho_mean, amp, dipoles, b->stream);
    BTRY(hipGetLastError());
    BTRY(hipStreamSynchronize(b->stream));
    b->nt = 0;
    b->aux_ok = false;
    return LB3D_OK;
}

int lb3d_gpu_batch_upload_sites(lb3d_batch *b, int32_t k, const double *whole_N, int32_t halo, int32_t sd)
{
    clear_error();
    if (!b || k < 0 || k >= b->nlat || !whole_N) return set_error(LB3D_EINVAL, "bad argument");
    std::vector<unsigned char> fb, dbb;
    int rc = aos_to_device_layout(b->p[k], whole_N, halo, sd, fb, dbb);
    if (rc) return rc;
    DevBufs v = view(b, k);
    BTRY(hipMemcpyAsync(v.f[b->cur], fb.data(), fb.size(), hipMemcpyHostToDevice, b->stream));
    if (b->C == 3) BTRY(hipMemcpyAsync(v.db, dbb.data(), dbb.size(), hipMemcpyHostToDevice, b->stream));
    BTRY(hipStreamSynchronize(b->stream));
    b->nt = 0;
    b->aux_ok = false;
    return LB3D_OK;
}

int lb3d_gpu_batch_get_populations(lb3d_batch *b, int32_t k, double *out)
{
    clear_error();
    if (!b || k < 0 || k >= b->nlat || !out) return set_error(LB3D_EINVAL, "bad argument");
    size_t bytes = (size_t)b->C * 19 * b->ns[k] * b->rs;
    std::vector<unsigned char> h(bytes);
    DevBufs v = view(b, k);
    BTRY(hipMemcpyAsync(h.data(), v.f[b->cur], bytes, hipMemcpyDeviceToHost, b->stream));
    BTRY(hipStreamSynchronize(b->stream));
    device_layout_to_pops(b->p[k], h.data(), out);
    return LB3D_OK;
}

int lb3d_gpu_batch_run_steps(lb3d_batch *b, int32_t nsteps)
{
    clear_error();
    if (!b || nsteps < 0) return set_error(LB3D_EINVAL, "bad argument");
    if (nsteps == 0) return LB3D_OK;
    int rc = ensure_graph(b);
    if (rc) return rc;
    if (b->cm_cols && !b->aux_ok) prime_aux(b);
    int left = nsteps;
    if (b->cur == 1) {
        enqueue_batch_step(b);
        left--;
    }
    for (; left >= 2; left -= 2) BTRY(hipGraphLaunch(b->graph, b->stream));
    if (left == 1) enqueue_batch_step(b);
    BTRY(hipGetLastError());
    b->nt += nsteps;
    b->aux_ok = b->cm_cols != 0;
    return LB3D_OK;
}

int lb3d_gpu_batch_run_steps_timed(lb3d_batch *b, int32_t nsteps, double *ms)
{
    clear_error();
    if (!b || !ms || nsteps < 1) return set_error(LB3D_EINVAL, "bad argument");
    // prepared before the first event: the graph, and the moments the
    // collide-moments plan carries
    int rc = ensure_graph(b);
    if (rc) return rc;
    if (b->cm_cols && !b->aux_ok) prime_aux(b);
    hipEvent_t e[2] = {nullptr, nullptr};
    struct Ev {
        hipEvent_t *e;
        ~Ev()
        {
            for (int j = 0; j < 2; j++)
                if (e[j]) (void)hipEventDestroy(e[j]);
        }
    } guard{e};
    BTRY(hipEventCreate(&e[0]));
    BTRY(hipEventCreate(&e[1]));
    BTRY(hipEventRecord(e[0], b->stream));
    if ((rc = lb3d_gpu_batch_run_steps(b, nsteps))) return rc;
    BTRY(hipEventRecord(e[1], b->stream));
    BTRY(hipEventSynchronize(e[1]));
    float t = 0.f;
    BTRY(hipEventElapsedTime(&t, e[0], e[1]));
    *ms = t;
    return LB3D_OK;
}

int lb3d_gpu_batch_synchronize(lb3d_batch *b)
{
    clear_error();
    if (!b) return set_error(LB3D_EINVAL, "NULL batch");
    BTRY(hipStreamSynchronize(b->stream));
    return LB3D_OK;
}

int lb3d_gpu_batch_diagnostics(lb3d_batch *b, int32_t k, double mass[3], double momentum[3], int64_t counts[3])
{
    clear_error();
    if (!b || k < 0 || k >= b->nlat || !mass || !momentum || !counts) return set_error(LB3D_EINVAL, "bad argument");
    b->L[k]->diagnostics(view(b, k), b->partials, b->stream);
    BTRY(hipGetLastError());
    unsigned nb = b->L[k]->diag_blocks();
    std::vector<double> h((size_t)nb * NDIAG);
    BTRY(hipMemcpyAsync(h.data(), b->partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    BTRY(hipStreamSynchronize(b->stream));
    double acc[NDIAG] = {0};
    for (unsigned q = 0; q < nb; q++)
        for (int j = 0; j < NDIAG; j++) acc[j] = acc[j] + h[(size_t)q * NDIAG + j];
    for (int j = 0; j < 3; j++) {
        mass[j] = j < b->C ? acc[j] : 0.0;
        momentum[j] = acc[3 + j];
        counts[j] = (int64_t)llround(acc[6 + j]);
    }
    return LB3D_OK;
}

int lb3d_gpu_batch_sanity(lb3d_batch *b, int32_t k, lb3d_sanity *out)
{
    clear_error();
    if (!b || k < 0 || k >= b->nlat || !out) return set_error(LB3D_EINVAL, "bad argument");
    b->L[k]->sanity(view(b, k), b->partials, b->stream);
    BTRY(hipGetLastError());
    unsigned nb = b->L[k]->diag_blocks();
    std::vector<double> h((size_t)nb * NSAN);
    BTRY(hipMemcpyAsync(h.data(), b->partials, h.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    BTRY(hipStreamSynchronize(b->stream));
    const lb3d_params &p = b->p[k];
    reduce_sanity(h.data(), nb, p.ncomp, p.nx, p.ny, p.nz, out);
    return LB3D_OK;
}

int lb3d_gpu_batch_set_plan(lb3d_batch *b, int32_t sequential, int32_t nwg, int32_t xcd_group)
{
    clear_error();
    if (!b) return set_error(LB3D_EINVAL, "NULL batch");
    if (!sequential && nwg < 1) return set_error(LB3D_EINVAL, "nwg must be >= 1");
    if (xcd_group < 0) return set_error(LB3D_EINVAL, "xcd_group must be >= 0");
    b->sequential = sequential != 0;
    b->xg = (uint32_t)xcd_group;
    b->B->xg = b->xg;
    if (!b->sequential) b->nwg = xcd_round((unsigned)nwg < b->nitems ? (unsigned)nwg : b->nitems);
    b->cm_cols = 0;
    return drop_graph(b);
}

int lb3d_gpu_batch_set_collide_moments(lb3d_batch *b, int32_t tile_cols, int32_t tile_threads, int32_t lz,
                                       int32_t xcd_group)
{
    clear_error();
    if (!b) return set_error(LB3D_EINVAL, "NULL batch");
    if (xcd_group < 0) return set_error(LB3D_EINVAL, "xcd_group must be >= 0");
    drop_graph(b);
    if (tile_cols == 0) {
        b->cm_cols = 0;
        return LB3D_OK;
    }
    if (int rc = build_cm(b, tile_cols, lz, tile_threads)) return rc;
    b->xg = (uint32_t)xcd_group;
    b->B->xg = b->xg;
    return LB3D_OK;
}

int lb3d_gpu_batch_set_groups(lb3d_batch *b, uint64_t group_bytes)
{
    clear_error();
    if (!b) return set_error(LB3D_EINVAL, "NULL batch");
    b->group_bytes = group_bytes;
    make_groups(b);
    return drop_graph(b);
}

// Times the persistent launches over grid sizes, and the lattice-by-lattice
// sequence, on the batch's own state; trial steps write f[1-cur] without
// swapping (the ternary dipole field is saved and restored).
int lb3d_gpu_batch_autotune(lb3d_batch *b, int32_t trial_steps, char *report, int32_t report_len)
{
    clear_error();
    if (!b) return set_error(LB3D_EINVAL, "NULL batch");
    if (trial_steps < 1) trial_steps = 1;
    struct Cand {
        bool seq;
        unsigned nwg;
        uint32_t xg;
        uint64_t gb;     // group bytes of the grouped persistent plan (0: one group)
        int cols = 0;    // collide-moments plan: tile columns (0: off), z-chunk, work-group size
        int lz = 0;
        int thr = 256;
    };
    std::vector<Cand> cands = {{true, 0, 0, 0}};
    for (unsigned g : {256u, 512u, 1024u, 2048u, 4096u, 8192u, 16384u})
        if (g < b->nitems) cands.push_back({false, g, 0, 0});
    cands.push_back({false, xcd_round(b->nitems), 0, 0});
    if (b->B->cm_supported(b->C))
        for (int thr : {256, 512})
            for (int cols : {30, 46, 62, 94})
                for (int lz : {8, 16})
                    if (cols <= thr - 2) cands.push_back({false, 0, 16, 0, cols, lz, thr});
    // the trials overwrite the dipole fields (the dipole pass writes them in
    // place); the guard restores them and the plan if a trial fails, and
    // frees the copy and the events on every path
    struct Guard {
        lb3d_batch *b;
        bool seq;
        unsigned nwg;
        uint32_t xg;
        uint64_t gb;
        int cols, lz, thr;
        size_t db_bytes = 0;
        void *db_save = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        bool restore = true;
        ~Guard()
        {
            if (restore) {
                b->sequential = seq;
                b->nwg = nwg;
                b->B->xg = xg;
                b->group_bytes = gb;
                make_groups(b);
                if (cols) {
                    if (build_cm(b, cols, lz, thr) == LB3D_OK) b->B->xg = xg;
                } else {
                    b->cm_cols = 0;
                }
                if (db_save) {
                    (void)hipMemcpyAsync(b->db, db_save, db_bytes, hipMemcpyDeviceToDevice, b->stream);
                    (void)hipStreamSynchronize(b->stream);
                }
            }
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (db_save) (void)hipFree(db_save);
        }
    } g{b, b->sequential, b->nwg, b->xg, b->group_bytes, b->cm_cols, b->cm_lz, b->cm_threads};
    // the trials overwrite aux(0), the scratch of the other plans
    b->aux_ok = false;
    g.db_bytes = b->C == 3 ? (b->db_off.back() + 3 * b->ns.back()) * b->rs : 0;
    if (g.db_bytes) {
        BTRY(hipMalloc(&g.db_save, g.db_bytes));
        BTRY(hipMemcpyAsync(g.db_save, b->db, g.db_bytes, hipMemcpyDeviceToDevice, b->stream));
    }
    BTRY(hipEventCreate(&g.e0));
    BTRY(hipEventCreate(&g.e1));
    hipEvent_t e0 = g.e0, e1 = g.e1;
    double best = 1e300;
    Cand bc = {b->sequential, b->nwg, b->xg, b->group_bytes, b->cm_cols, b->cm_lz, b->cm_threads};
    std::string rep;
    auto time_cand = [&](const Cand &c) -> int {
        b->sequential = c.seq;
        b->nwg = c.nwg;
        b->B->xg = c.xg;
        b->group_bytes = c.gb;
        make_groups(b);
        if (c.cols) {
            if (c.cols != b->cm_cols || c.lz != b->cm_lz || c.thr != b->cm_threads)
                if (int rc = build_cm(b, c.cols, c.lz, c.thr)) return rc;
            prime_aux(b);
        } else {
            b->cm_cols = 0;
        }
        auto trial = [&]() {
            enqueue_batch_step(b);
            b->cur = 1 - b->cur;  // undo the swap: the state stays in f[cur]
        };
        trial();
        BTRY(hipEventRecord(e0, b->stream));
        for (int t = 0; t < trial_steps; t++) trial();
        BTRY(hipEventRecord(e1, b->stream));
        BTRY(hipGetLastError());
        BTRY(hipEventSynchronize(e1));
        float ms = 0.f;
        BTRY(hipEventElapsedTime(&ms, e0, e1));
        double per = ms / trial_steps;
        char buf[128];
        if (c.cols)
            std::snprintf(buf, sizeof buf, "collide-moments cols=%d threads=%d lz=%d xg=%u %.4f ms\n", c.cols, c.thr,
                          c.lz, c.xg, per);
        else
            std::snprintf(buf, sizeof buf, "%s nwg=%u xg=%u groups=%zu (%llu MiB) %.4f ms\n",
                          c.seq ? "sequential" : "persistent", c.nwg, c.xg, c.seq ? (size_t)0 : b->groups.size() - 1,
                          (unsigned long long)(c.gb >> 20), per);
        rep += buf;
        if (per < best) {
            best = per;
            bc = c;
        }
        return LB3D_OK;
    };
    for (const Cand &c : cands)
        if (int rc = time_cand(c)) return rc;
    if (!bc.seq) {  // the XCD chunking of the best persistent grid or collide-moments plan
        const Cand top = bc;
        for (uint32_t xg : {1u, 4u, 16u, 64u, 256u})
            if (xg != top.xg)
                if (int rc = time_cand({false, top.nwg, xg, 0, top.cols, top.lz, top.thr})) return rc;
    }
    if (!bc.seq && !bc.cols) {  // then groups of lattices sized for the memory-side cache
        const Cand top = bc;
        uint64_t total = 0;
        for (int k = 0; k < b->nlat; k++) total += (uint64_t)b->C * 19u * b->ns[k] * b->rs;
        // measured on the 48-lattice ensemble: 192-320 MiB groups 0.3-1.9%
        // faster per step than one group, depending on where the group
        // boundaries fall (hence the 32 MiB steps); 160 MiB and smaller lose
        // more to their launches than they gain, 336 MiB and larger evict
        // what the collide re-reads
        for (uint64_t mib : {192u, 224u, 256u, 288u, 320u})
            if ((mib << 20) < total)
                if (int rc = time_cand({false, top.nwg, top.xg, mib << 20})) return rc;
    }
    if (!bc.seq && !bc.cols && bc.gb) {  // the XCD chunk again, with the chosen groups
        const Cand top = bc;
        for (uint32_t xg : {4u, 16u, 64u, 256u})
            if (xg != top.xg)
                if (int rc = time_cand({false, top.nwg, xg, top.gb})) return rc;
    }
    if (g.db_save) {
        BTRY(hipMemcpyAsync(b->db, g.db_save, g.db_bytes, hipMemcpyDeviceToDevice, b->stream));
        BTRY(hipStreamSynchronize(b->stream));
    }
    g.restore = false;
    b->sequential = bc.seq;
    b->nwg = bc.nwg;
    b->xg = bc.xg;
    b->B->xg = bc.xg;
    b->group_bytes = bc.gb;
    make_groups(b);
    if (bc.cols) {
        if (bc.cols != b->cm_cols || bc.lz != b->cm_lz || bc.thr != b->cm_threads)
            if (int rc = build_cm(b, bc.cols, bc.lz, bc.thr)) return rc;
    } else {
        b->cm_cols = 0;
    }
    b->aux_ok = false;
    drop_graph(b);
    char buf[128];
    if (bc.cols)
        std::snprintf(buf, sizeof buf, "chosen: collide-moments cols=%d threads=%d lz=%d xg=%u %.4f ms\n", bc.cols,
                      bc.thr, bc.lz, bc.xg, best);
    else
        std::snprintf(buf, sizeof buf, "chosen: %s nwg=%u xg=%u groups=%zu (%llu MiB) %.4f ms\n",
                      bc.seq ? "sequential" : "persistent", bc.nwg, bc.xg, bc.seq ? (size_t)0 : b->groups.size() - 1,
                      (unsigned long long)(bc.gb >> 20), best);
    rep += buf;
    if (report && report_len > 0) {
        std::strncpy(report, rep.c_str(), (size_t)report_len - 1);
        report[report_len - 1] = 0;
    }
    return LB3D_OK;
}

// Time per step spent in each of the three phases, summed over the groups
// of the persistent plan (or the lattices of the sequential plan) in the
// order the step runs them, over nsteps timesteps: HIP events on the batch's
// stream after every launch.
int lb3d_gpu_batch_profile_steps(lb3d_batch *b, int32_t nsteps, double ms_phase[3], double *ms_per_step)
{
    clear_error();
    if (!b || nsteps < 1 || !ms_phase || !ms_per_step) return set_error(LB3D_EINVAL, "bad argument");
    const bool cm = b->cm_cols != 0;
    if (cm && !b->aux_ok) prime_aux(b);
    const size_t units = cm ? 1 : (b->sequential ? (size_t)b->nlat : b->groups.size() - 1);
    const size_t per = 1 + 3 * units;  // events per step
    // destroyed when the scope ends, also on an early error return
    struct Events {
        std::vector<hipEvent_t> e;
        ~Events()
        {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } evs{std::vector<hipEvent_t>((size_t)nsteps * per, nullptr)};
    std::vector<hipEvent_t> &ev = evs.e;
    for (auto &e : ev) BTRY(hipEventCreate(&e));
    for (int st = 0; st < nsteps; st++) {
        size_t q = (size_t)st * per;
        BTRY(hipEventRecord(ev[q++], b->stream));
        for (size_t u = 0; u < units; u++) {
            for (int ph = 0; ph < 3; ph++) {
                if (cm) {
                    // the plan's order: dipole, collide, edge moments; their
                    // times go to ms_phase[1], [2] and [0]
                    cm_phase(b, (ph + 1) % 3);
                } else if (b->sequential) {
                    DevBufs v = view(b, (int)u);
                    if (ph == 0) b->L[u]->moments(v, b->stream);
                    if (ph == 1) b->L[u]->dipole(v, b->stream);
                    if (ph == 2) b->L[u]->collide(v, b->stream);
                } else {
                    group_phase(b, u, ph);
                }
                BTRY(hipEventRecord(ev[q++], b->stream));
            }
        }
        b->cur = 1 - b->cur;
    }
    BTRY(hipGetLastError());
    BTRY(hipStreamSynchronize(b->stream));
    double tot[3] = {0, 0, 0}, all = 0;
    for (int st = 0; st < nsteps; st++) {
        for (size_t j = 0; j + 1 < per; j++) {
            float m = 0.f;
            BTRY(hipEventElapsedTime(&m, ev[(size_t)st * per + j], ev[(size_t)st * per + j + 1]));
            tot[cm ? (j + 1) % 3 : j % 3] += m;
            all += m;
        }
    }
    for (int ph = 0; ph < 3; ph++) ms_phase[ph] = tot[ph] / nsteps;
    *ms_per_step = all / nsteps;
    b->nt += nsteps;
    b->aux_ok = cm;
    return LB3D_OK;
}

int lb3d_gpu_batch_get_nt(lb3d_batch *b, int32_t *nt)
{
    if (!b || !nt) return set_error(LB3D_EINVAL, "bad argument");
    *nt = b->nt;
    return LB3D_OK;
}

}  // extern "C"
