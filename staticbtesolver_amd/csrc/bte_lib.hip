// bte_lib.hip -- host side of the C ABI (include/bte.h): device-resident state, sweep schedule,
// outer source iteration with one hipGraph per iteration parity.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bte.h"
#include "bte_internal.hpp"
#include "bte_kernels.hpp"
#include "bte_bands.hpp"

using namespace bte;

// Reduction scratch of bte_solve_noalloc, in this library's code object: one copy of the array per
// device, 1 MiB.  Host code reaches a device's copy only through the address hipGetSymbolAddress
// returns for that device (g_scratch), never through the array's name.
constexpr size_t kScratchDoubles = size_t(1) << 17;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

// Slices of the scratch (offsets in doubles, multiples of 32 = 256 bytes):
//   margin partial sums of the band-temperature pass (one per work-group, <= 65536 by grid_for),
//   the margin, the partial sums (2 x 32) and totals (4) of launch_norms2, the Anderson dot-product partials
//   ((2m+1) x 64, m <= 64) and dots, and the heat flux per (BC, band) before its copy to the host.
constexpr size_t kScrPartial = 0, kScrMargin = 65536, kScrRedPart = kScrMargin + 32, kScrRedOut = kScrRedPart + 64,
                 kScrAaPart = kScrRedOut + 32, kScrAaDots = kScrAaPart + 129 * 64, kScrFlux = kScrAaDots + 160,
                 kScrFluxMax = kScratchDoubles - kScrFlux;

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};
hipError_t g_scratch_rc[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// Address of `device`'s scratch copy, resolved once per device with that device current.
double* scratch_of(int device, hipError_t* rc) {
    if (device < 0 || device >= kMaxDevices) {
        *rc = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_done[device]) {
        void* ptr = nullptr;
        g_scratch_rc[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        g_scratch[device] = g_scratch_rc[device] == hipSuccess ? static_cast<double*>(ptr) : nullptr;
        g_scratch_done[device] = true;
    }
    *rc = g_scratch_rc[device];
    return g_scratch[device];
}

// Kinds timed with opts.timing: "rhs" covers the right-hand sides and the diffuse re-emission
// before the inner solve, "solve" the inner-solve kernel alone (the sweep, or BiCGSTAB).
const char* kKernelNames[] = {"assemble", "sweep_schedule", "rhs", "solve", "temperature", "margin", "flux"};
enum { KA_ASM = 0, KA_SCHED, KA_RHS, KA_SOLVE, KA_TEMP, KA_MARGIN, KA_FLUX, KA_COUNT };
std::atomic<unsigned long long> g_geo_tags{0};

}  // namespace

// Cholesky factor R (upper, R^T R = H + lam I) of the Gram matrix H of the Anderson history
// columns, in the order the history keeps them (oldest first), carried from step to step.
struct AaChol {
    static constexpr int M = bte::kAaMaxCols;
    int n = -1;  // columns factored; -1: none valid
    double lam = 0;
    double R[M * M];  // R[i * M + j], j >= i
};

struct bte_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // problem (host copies)
    int dim = 0, N = 0, F = 0, D = 0, B = 0, DM = 0, num_theta = 0, num_phi = 0;
    double solid_angle = 4 * kPi, T_ref = 0;
    std::vector<double> h_vol, h_area, h_normal, h_omega, h_dir, h_S, h_v, h_tau, h_C, h_Lr, h_bcT;
    std::vector<int> h_nbr, h_bci, h_bct;
    // Cells are stored in an internal order chosen for memory locality (locality_order); cell_of[i]
    // is the caller's index of internal cell i (empty: the caller's order).  Every per-cell array
    // the library holds, uploads or computes is in internal order; the entry points map the
    // per-cell arrays they take and return.
    std::vector<int> cell_of;
    bool ordered = false;  // locality_order has run for the current geometry
    bool have_geo = false, have_quad = false, have_bands = false, have_bcs = false;
    bool prepared = false;
    // device: geometry / quadrature / bands
    DBuf vol, area, normal, nbr, omega, dir, S, w_eff, bv, btau, bC, bLr, vt;
    // boundary
    DBuf bf_cell, bf_face, bfid, mirror, ediff;
    int n_bf = 0;
    bool has_diffuse = false;
    // assembly + sweep schedule
    int W = 0;
    DBuf dg, ell_col, ell_val, overflow, dg32, ell_val32;
    DBuf lvl_off, lvl_cnt, lvl_ptr, perm, swp_col, swp_val, swp_dg, swp_val32, swp_dg32, swp_vw, swp_bnd, swp_pos;
    DBuf lvl_sync, dir_fb;             // pipelined sweep: synchronisation levels, directions with HBM reads
    DBuf swp_pki, swp_pkd;             // pipelined sweep: packed level-order operand records
    int ring_r = 0;                    // ring size those were computed for (ring_size(ctx), 0: none)
    int ring_r2 = 0;                   // the wide ring's size (ring_wide(ctx), 0: none), flags in bits 2, 3
    int max_levels = 0, lagged = 0;
    int ring_span[7] = {};             // max positions spanned by 64, 32, ..., 2, 1 consecutive sweep levels
    int max_back = 0;                  // max over cells of (end of its level) - (upwind position)
    std::vector<int> h_lvl_cnt;  // sweep levels per direction
    DBuf bnd_list;               // level-order entries d*N+i of cells with an incoming boundary face
    int n_bnd_list = 0;
    int max_lds = 65536;  // dynamic LDS bytes a work-group may use on this device
    bool have_sweep = false, have_fp32 = false;
    // state + work
    DBuf eeA, eeB, dens, T, Tband, partial, margin, bflux;
    // source temperatures: tsrc[c] is the temperature the emission density of cell c was formed
    // from (every writer of dens stores it), tlo the same in level order per direction; tsrc_ok
    // while that holds (a density uploaded as such breaks it until the next bte_begin)
    DBuf tsrc, tlo;
    bool tsrc_ok = false;
    DBuf wb, wr64, wr, wrh, wp, wv, wph, wsh, wt, wy, wf, iters, fails;
    double* ee_final = nullptr;  // which of eeA/eeB holds the last iterate
    // layout of the intensity arrays (band-major, or interleaved for a pipelined 3D sweep: set by
    // bte_begin, band-major whenever work buffers are (re)sized)
    EeIdx eei{};
    bool ee_inter = false;
    bool run_inter = false;  // the layout of the last solve's intensities (ee_final)
    double* h_margin_pinned = nullptr;
    // coherent pinned host memory that kernels store into directly (no copy, read after a
    // synchronisation): [0] the margin of the last outer iteration, [64, 64 + 2 m + 1) the dot
    // products of the last Anderson step; d_sync is its device address
    double* h_sync = nullptr;
    double* d_sync = nullptr;
    long long sync_seq = 0;  // last value k_signal stored into h_sync[kSyncFlag] (wait_stream)
    // Anderson step launched behind the margin of the iteration (one synchronisation for both), and
    // the host mixing state from before it, restored when the margin ends the solve instead
    bool aa_spec = false, aa_spec_finish = false;
    int aa_saved_steps = 0;
    bool aa_saved_skip = false;
    std::vector<int> aa_saved_order;
    int aa_slot = 0;            // history column of the step between aa_launch and aa_finish
    int aa_ncols = 0;           // filled history columns [0, aa_ncols) of that step
    bool aa_skip_now = false;   // that step follows an accepted correction
    // graphs, and the options and launch shape they were captured with (graph_valid: they exist)
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    bool graph_valid = false;
    bte_opts graph_opts{};
    int graph_block = 0, graph_variant = 0;
    bool graph_inter = false;
    // stats
    double ms[KA_COUNT] = {0};
    long long calls[KA_COUNT] = {0};
    double spmv_bytes = 0;
    long long spmv_rows = 0;
    int block_used = 0, variant_used = 0, group_used = 0;
    // autotuned launch shape per solver form: work-group size and sweep variant (0 = not tuned)
    int tuned_block[2][2][2][4] = {};  // [precision][precond][direct][requested band group: auto, 1, 2, 4]
    // iteration state between bte_begin and bte_finish
    bte_opts run_opts{};
    int run_shape_block = 256, run_shape_variant = 2;
    bool run_graph = false, run_active = false, run_converged = false;
    int run_iter = 0, run_accepted = 0, run_proposed = 0;
    std::vector<double> run_margins;
    double run_solve_seconds = 0, run_setup_seconds = 0;
    // Anderson mixing state (ring of m difference columns)
    DBuf aa_x, aa_xprev, aa_fprev, aa_f, aa_dX, aa_dF, aa_dG, aa_part, aa_dots, aa_g, aa_bcells, aa_sc;
    int aa_m = 0, aa_nb = 0;
    long long aa_Z = 0;  // state length N + S * nb
    bool aa_first = true;
    std::vector<int> aa_order;  // active slots, oldest first
    std::vector<double> aa_H;   // Gram matrix of the dF columns by slot [m*m]
    AaChol aa_chol;             // its factor in aa_order's order
    bool aa_popped = false;     // the last aa_launch dropped the oldest column
    double aa_best = 0;
    double aa_last_ff = 0;      // squared residual norm of the last evaluated Anderson state
    int aa_steps = 0;
    unsigned long long geo_tag = 0;  // changes with every bte_set_geometry
    bool aa_plain = false;      // state bookkeeping without mixing (plain iteration + learned correction)
    bool aa_skip = false;       // the next state step follows an accepted correction
    // learned correction
    bte_gnn* gnn = nullptr;     // attached network (not owned)
    DBuf learn_stat, learn_backup, learn_X, red_part, red_out;
    bool learn_stat_ok = false, learn_pending = false;
    bool aa_cells_ok = false;  // aa_bcells / aa_sc / aa_nb hold this geometry's, BC set's and band set's
    double learn_Tmid = 0, learn_span = 1, learn_ref_ff = 0, learn_ratio = 0, learn_jump = 0;
    bool judge_spec = false;  // the pending correction's judge norms were launched behind its iteration
    int learn_streak = 0;       // consecutive rejections
    int run_rejected = 0;
    int tuned_variant[2][2][2][4] = {};  // sweep variant (direct form) or band group (Krylov forms)
    // bte_solve_noalloc: this device's scratch copy (null if it could not be resolved) and whether
    // the reductions use it for the current call
    double* scratch_dev = nullptr;
    hipError_t scratch_rc = hipSuccess;
    bool use_scratch = false;
};

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);                         \
            return BTE_ERR_HIP;                                                                   \
        }                                                                                         \
    } while (0)

#define SETERR(code, msg)   \
    do {                    \
        ctx->err = (msg);   \
        return (code);      \
    } while (0)

namespace {

template <class T>
hipError_t upload(DBuf& d, const std::vector<T>& h, hipStream_t s) {
    hipError_t e = d.ensure(h.size() * sizeof(T));
    if (e != hipSuccess) return e;
    if (h.empty()) return hipSuccess;
    return hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

int grid_for(long long n, int block) {
    long long g = (n + block - 1) / block;
    return (int)std::max<long long>(1, std::min<long long>(g, 65536));
}

// Work-groups of the diffuse re-emission pass (one thread per face and group of diffuse_group bands
// of the intensity layout in use).
int diffuse_grid(const bte_ctx* c) {
    const int G = diffuse_group(c->eei);
    return grid_for((long long)((c->B + G - 1) / G) * c->n_bf, 256);
}

// Work-groups (64 threads) of the boundary right-hand-side pass: one per chunk of 64 list entries
// and band, a multiple of 8 up to the grid cap.
// List entries per work-group of k_rhs_levels_bnd (3D 200k: 597 us with 64, 581 with 128, 547 with
// 256); the batch's multi-problem launch keeps 64.
constexpr int kBndEntries = 256;
int rhs_bnd_grid(const bte_ctx* c, int entries = 64) {
    return (int)std::min<long long>(rhs_bnd_units(c->n_bnd_list, c->B, entries), 65536);
}

Geo geo_of(bte_ctx* c) {
    return Geo{c->dim, c->N, c->F, c->vol.as<double>(), c->area.as<double>(), c->normal.as<double>(),
               c->nbr.as<int>()};
}
Quad quad_of(bte_ctx* c) {
    return Quad{c->D, c->omega.as<double>(), c->dir.as<double>(), c->S.as<double>(), c->eei};
}

// Intensity layout of the next launches: band-major, or direction-major with each cell's bands
// interleaved (EeIdx).
void set_layout(bte_ctx* c, bool interleaved) {
    c->ee_inter = interleaved;
    c->eei = interleaved ? ee_interleaved(c->B, c->N) : ee_band_major(c->D, c->N);
}
Bands bands_of(bte_ctx* c) {
    return Bands{c->B, c->bv.as<double>(), c->btau.as<double>(), c->bC.as<double>(), c->bLr.as<double>(),
                 c->solid_angle, c->T_ref};
}
Bnd bnd_of(bte_ctx* c) {
    Bnd b{};
    b.nbc = (int)c->h_bci.size();
    for (int k = 0; k < b.nbc; k++) {
        b.bci[k] = c->h_bci[k];
        b.bct[k] = c->h_bct[k];
        b.bcT[k] = c->h_bcT[k];
    }
    b.n_bf = c->n_bf;
    b.bf_cell = c->bf_cell.as<int>();
    b.bf_face = c->bf_face.as<int>();
    b.bfid = c->bfid.as<int>();
    b.mirror = c->mirror.as<int>();
    b.ediff = c->ediff.as<double>();
    b.has_diffuse = c->has_diffuse ? 1 : 0;
    return b;
}

// Reduction buffers: the context's own, or the device's scratch slices during bte_solve_noalloc.
double* partial_of(bte_ctx* c) { return c->use_scratch ? c->scratch_dev + kScrPartial : c->partial.as<double>(); }
double* margin_of(bte_ctx* c) { return c->use_scratch ? c->scratch_dev + kScrMargin : c->margin.as<double>(); }
double* red_part_of(bte_ctx* c) { return c->use_scratch ? c->scratch_dev + kScrRedPart : c->red_part.as<double>(); }
double* red_out_of(bte_ctx* c) { return c->use_scratch ? c->scratch_dev + kScrRedOut : c->red_out.as<double>(); }
double* aa_part_of(bte_ctx* c) { return c->use_scratch ? c->scratch_dev + kScrAaPart : c->aa_part.as<double>(); }
double* aa_dots_of(bte_ctx* c) { return c->use_scratch ? c->scratch_dev + kScrAaDots : c->aa_dots.as<double>(); }
double* bflux_of(bte_ctx* c) { return c->use_scratch ? c->scratch_dev + kScrFlux : c->bflux.as<double>(); }

// Mirror direction of the specular types, StaticBTESolver.cpp:205-252 (-1 where the reference exits).
int mirror_of(int dim, int num_phi, int type, int d) {
    const int p = num_phi;
    const int it = d / (4 * p), ip = d % (4 * p);
    if (type == 31) {
        if (dim == 2) return it * 4 * p + ((ip >= 2 * p) ? 6 * p - ip - 1 : 2 * p - ip - 1);
        if (dim == 3) return it * 4 * p + (4 * p - ip - 1);
        return -1;
    }
    if (type == 32) {
        if (dim == 2) return it * 4 * p + (4 * p - ip - 1);
        if (dim == 3) return it * 4 * p + (2 * p - (ip % (2 * p)) - 1) + (ip / (2 * p)) * 2 * p;
        return -1;
    }
    if (type == 33) {
        if (dim < 3) return -1;
        return ip + (2 * p - it - 1) * 4 * p;  // theta index mirrored with num_phi (quirk of :248-251)
    }
    return -1;
}

// Downwind level schedule of one direction from the ELL columns (Kahn's algorithm, level by level).
// Edges whose source is still unplaced when no cell is ready are lagged (dropped from the sweep).
// Each level lists its cells in increasing index, so the lanes of a wave that write a level's
// intensities store to increasing addresses, and cells adjacent in memory share store transactions.
void level_schedule(int N, int W, const int* col /* [W*N] of this direction */, std::vector<int>& order,
                    std::vector<int>& lptr, std::vector<int>& scol, int& lagged) {
    std::vector<int> indeg(N, 0), outcnt(N + 1, 0);
    for (int k = 0; k < W; k++)
        for (int c = 0; c < N; c++) {
            const int j = col[(size_t)k * N + c];
            if (j >= 0) {
                indeg[c]++;
                outcnt[j + 1]++;
            }
        }
    for (int i = 0; i < N; i++) outcnt[i + 1] += outcnt[i];
    std::vector<int> out(outcnt[N]), fill(outcnt.begin(), outcnt.end() - 1);
    for (int k = 0; k < W; k++)
        for (int c = 0; c < N; c++) {
            const int j = col[(size_t)k * N + c];
            if (j >= 0) out[fill[j]++] = c;
        }
    std::vector<char> placed(N, 0), lag((size_t)W * N, 0);
    order.clear();
    order.reserve(N);
    lptr.assign(1, 0);
    std::vector<int> frontier, next;
    for (int c = 0; c < N; c++)
        if (indeg[c] == 0) frontier.push_back(c);
    int scan = 0;
    lagged = 0;
    while ((int)order.size() < N) {
        if (frontier.empty()) {
            while (placed[scan]) scan++;
            const int pick = scan;
            for (int k = 0; k < W; k++) {
                const int j = col[(size_t)k * N + pick];
                if (j >= 0 && !placed[j]) {
                    lag[(size_t)k * N + pick] = 1;
                    lagged++;
                }
            }
            frontier.push_back(pick);
        }
        std::sort(frontier.begin(), frontier.end());
        for (int c : frontier) {
            placed[c] = 1;
            order.push_back(c);
        }
        lptr.push_back((int)order.size());
        next.clear();
        for (int c : frontier)
            for (int e = outcnt[c]; e < outcnt[c + 1]; e++) {
                const int o = out[e];
                if (!placed[o] && --indeg[o] == 0) next.push_back(o);
            }
        frontier.swap(next);
    }
    scol.assign((size_t)W * N, -1);
    for (int i = 0; i < N; i++) {
        const int c = order[i];
        for (int k = 0; k < W; k++)
            scol[(size_t)k * N + i] = lag[(size_t)k * N + c] ? -1 : col[(size_t)k * N + c];
    }
}

struct Timer {
    bte_ctx* ctx;
    int kind;
    hipEvent_t a = nullptr, b = nullptr;
    bool on;
    Timer(bte_ctx* c, int k, bool enabled) : ctx(c), kind(k), on(enabled) {
        if (on) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, ctx->stream);
        }
    }
    ~Timer() {
        if (on) {
            (void)hipEventRecord(b, ctx->stream);
            (void)hipEventSynchronize(b);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, a, b);
            ctx->ms[kind] += ms;
            ctx->calls[kind]++;
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    }
};

void drop_graphs(bte_ctx* ctx) {
    for (auto& g : ctx->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    ctx->graph_valid = false;
    ctx->run_active = false;  // captured launches referenced buffers that may change
}

}  // namespace

extern "C" {

void bte_default_opts(bte_opts* o) {
    std::memset(o, 0, sizeof(*o));
    o->precision = BTE_PREC_FP64;
    o->precond = BTE_PC_SWEEP;
    o->max_iter = 10000;
    o->stop_margin = 1e-5;
    o->inner_tol = 1e-9;
    o->inner_maxit = 1000;
    o->refine_steps = 4;
    o->block = 0;
    o->use_graph = 1;
    o->verbose = 0;
    o->timing = 0;
    o->learned = 0;
    o->learned_every = 8;
    o->direct_sweep = 1;
    o->sweep_variant = 0;
    o->accel = BTE_ACCEL_NONE;
    o->accel_memory = 20;
    o->band_group = 0;
    o->inner_tol32 = 1e-5;
    o->intensity_layout = 0;
}

int bte_ctx_create(int device, bte_ctx** out) {
    if (!out) return BTE_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return BTE_ERR_NODEVICE;
    bte_ctx* ctx = new bte_ctx();
    if (device < 0) (void)hipGetDevice(&device);
    if (device >= n) {
        delete ctx;
        return BTE_ERR_ARG;
    }
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_margin_pinned, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&ctx->h_sync, 256 * sizeof(double), hipHostMallocCoherent | hipHostMallocMapped) !=
            hipSuccess ||
        hipHostGetDevicePointer((void**)&ctx->d_sync, ctx->h_sync, 0) != hipSuccess) {
        delete ctx;
        return BTE_ERR_HIP;
    }
    std::memset(ctx->h_sync, 0, 256 * sizeof(double));  // no sequence number stored yet
    // this device's copy of the scratch array, with the device current (bte_solve_noalloc)
    ctx->scratch_dev = scratch_of(device, &ctx->scratch_rc);
    *out = ctx;
    return BTE_OK;
}

void bte_ctx_destroy(bte_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    drop_graphs(ctx);
    if (ctx->h_margin_pinned) (void)hipHostFree(ctx->h_margin_pinned);
    if (ctx->h_sync) (void)hipHostFree(ctx->h_sync);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* bte_last_error(const bte_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int bte_set_geometry(bte_ctx* ctx, int dim, int n_cell, int n_face, const double* volume, const double* face_area,
                     const double* face_normal, const int* neighbor) {
    if (!ctx) return BTE_ERR_ARG;
    if (dim < 1 || dim > 3 || n_face != dim + 1 || n_cell <= 0 || !volume || !face_area || !face_normal || !neighbor)
        SETERR(BTE_ERR_ARG, "bte_set_geometry: need dim in {1,2,3}, n_face = dim+1, n_cell > 0 and all arrays");
    HIPCHK(hipSetDevice(ctx->device));
    ctx->dim = dim;
    ctx->N = n_cell;
    ctx->F = n_face;
    const size_t NF = (size_t)n_cell * n_face;
    ctx->h_vol.assign(volume, volume + n_cell);
    ctx->h_area.assign(face_area, face_area + NF);
    ctx->h_normal.assign(face_normal, face_normal + 3 * NF);
    ctx->h_nbr.assign(neighbor, neighbor + NF);
    for (size_t i = 0; i < NF; i++)
        if (ctx->h_nbr[i] >= n_cell) SETERR(BTE_ERR_MESH, "bte_set_geometry: neighbour index out of range");
    HIPCHK(upload(ctx->vol, ctx->h_vol, ctx->stream));
    HIPCHK(upload(ctx->area, ctx->h_area, ctx->stream));
    HIPCHK(upload(ctx->normal, ctx->h_normal, ctx->stream));
    HIPCHK(upload(ctx->nbr, ctx->h_nbr, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->have_geo = true;
    ctx->prepared = false;
    ctx->ordered = false;
    ctx->cell_of.clear();
    ctx->learn_stat_ok = false;
    ctx->aa_cells_ok = false;
    ctx->geo_tag = ++g_geo_tags;
    drop_graphs(ctx);
    return BTE_OK;
}

int bte_set_quadrature(bte_ctx* ctx, int DM, int n_dir, const double* omega, const double* dir, const double* S,
                       int num_theta, int num_phi, double solid_angle, double T_ref) {
    if (!ctx) return BTE_ERR_ARG;
    if (n_dir <= 0 || !omega || !dir || !S || (DM != 2 && DM != 3) || num_phi <= 0)
        SETERR(BTE_ERR_ARG, "bte_set_quadrature: need n_dir > 0, DM in {2,3}, num_phi > 0 and all arrays");
    HIPCHK(hipSetDevice(ctx->device));
    ctx->DM = DM;
    ctx->D = n_dir;
    ctx->num_theta = num_theta;
    ctx->num_phi = num_phi;
    ctx->solid_angle = solid_angle;
    ctx->T_ref = T_ref;
    ctx->h_omega.assign(omega, omega + n_dir);
    ctx->h_dir.assign(dir, dir + 3 * n_dir);
    ctx->h_S.assign(S, S + 3 * n_dir);
    HIPCHK(upload(ctx->omega, ctx->h_omega, ctx->stream));
    HIPCHK(upload(ctx->dir, ctx->h_dir, ctx->stream));
    HIPCHK(upload(ctx->S, ctx->h_S, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->have_quad = true;
    ctx->prepared = false;
    drop_graphs(ctx);
    return BTE_OK;
}

int bte_set_bands(bte_ctx* ctx, int n_band, const double* v, const double* tau, const double* C, const double* Lr) {
    if (!ctx) return BTE_ERR_ARG;
    if (n_band <= 0 || !v || !tau || !C || !Lr) SETERR(BTE_ERR_ARG, "bte_set_bands: need n_band > 0 and all arrays");
    HIPCHK(hipSetDevice(ctx->device));
    ctx->B = n_band;
    ctx->h_v.assign(v, v + n_band);
    ctx->h_tau.assign(tau, tau + n_band);
    ctx->h_C.assign(C, C + n_band);
    ctx->h_Lr.assign(Lr, Lr + n_band);
    std::vector<double> vt(n_band);
    for (int b = 0; b < n_band; b++) vt[b] = v[b] * tau[b];
    HIPCHK(upload(ctx->bv, ctx->h_v, ctx->stream));
    HIPCHK(upload(ctx->btau, ctx->h_tau, ctx->stream));
    HIPCHK(upload(ctx->bC, ctx->h_C, ctx->stream));
    HIPCHK(upload(ctx->bLr, ctx->h_Lr, ctx->stream));
    HIPCHK(upload(ctx->vt, vt, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->have_bands = true;
    ctx->learn_stat_ok = false;
    ctx->aa_cells_ok = false;
    drop_graphs(ctx);
    return BTE_OK;
}

int bte_set_bcs(bte_ctx* ctx, int n_bc, const int* index, const int* type, const double* T) {
    if (!ctx) return BTE_ERR_ARG;
    if (n_bc < 0 || n_bc > kMaxBC || (n_bc > 0 && (!index || !type || !T)))
        SETERR(BTE_ERR_ARG, "bte_set_bcs: 0 <= n_bc <= 32 and all arrays");
    for (int k = 0; k < n_bc; k++)
        if (index[k] >= 0) SETERR(BTE_ERR_BC, "bte_set_bcs: boundary index must be -k-1 with k >= 0");
    ctx->h_bci.assign(index, index + n_bc);
    ctx->h_bct.assign(type, type + n_bc);
    ctx->h_bcT.assign(T, T + n_bc);
    ctx->have_bcs = true;
    ctx->prepared = false;
    ctx->learn_stat_ok = false;
    ctx->aa_cells_ok = false;
    drop_graphs(ctx);
    return BTE_OK;
}

}  // extern "C"

namespace {

// LDS ring size of the pipelined sweeps (G bands per work-group), 0 when they cannot run
int ring_size(const bte_ctx* ctx, int* sync_every = nullptr, int G = 1);
int ring_wide(const bte_ctx* ctx);

// Kahn level of every cell in the upwind graph of direction v (cell c waits for its neighbour
// across each face with v.n < 0); a cycle is broken at the lowest unplaced cell.
std::vector<int> upwind_levels(const bte_ctx* ctx, const double* v) {
    const int N = ctx->N, F = ctx->F;
    std::vector<int> indeg(N, 0), lvl(N, 0), start(N + 1, 0);
    std::vector<char> up((size_t)N * F, 0);
    for (int c = 0; c < N; c++)
        for (int f = 0; f < F; f++) {
            const size_t k = (size_t)c * F + f;
            const int nb = ctx->h_nbr[k];
            const double* n = ctx->h_normal.data() + 3 * k;
            if (nb >= 0 && v[0] * n[0] + v[1] * n[1] + v[2] * n[2] < 0) {
                up[k] = 1;
                indeg[c]++;
                start[nb + 1]++;
            }
        }
    for (int i = 0; i < N; i++) start[i + 1] += start[i];
    std::vector<int> down(start[N]), fill(start.begin(), start.end() - 1);
    for (int c = 0; c < N; c++)
        for (int f = 0; f < F; f++)
            if (up[(size_t)c * F + f]) down[fill[ctx->h_nbr[(size_t)c * F + f]]++] = c;
    std::vector<int> queue;
    queue.reserve(N);
    std::vector<char> placed(N, 0);
    for (int c = 0; c < N; c++)
        if (!indeg[c]) queue.push_back(c);
    int scan = 0;
    for (size_t h = 0; (int)queue.size() < N || h < queue.size();) {
        if (h == queue.size()) {
            while (placed[scan] || indeg[scan] == 0) scan++;
            indeg[scan] = 0;
            queue.push_back(scan);
        }
        const int c = queue[h++];
        placed[c] = 1;
        for (int e = start[c]; e < start[c + 1]; e++) {
            const int o = down[e];
            lvl[o] = std::max(lvl[o], lvl[c] + 1);
            if (--indeg[o] == 0) queue.push_back(o);
        }
    }
    return lvl;
}

// Internal cell order for memory locality.  The pipelined sweeps write each system's intensities
// cell by cell as the levels of its direction arrive, and the levels of different directions cut
// the mesh along different planes; an order in which cells close in space are close in memory
// lets the writes of one level share cache lines.  The cells' coordinates are not inputs, so the
// upwind levels of the quadrature directions closest to +x, +y (and +z) stand in for them (a cell's
// level grows with its distance from the upwind wall), and the cells are sorted along the Morton
// curve of those level coordinates.  Stored geometry (host and device) is permuted in place and
// the neighbour indices renumbered; cell_of keeps the caller's index of each internal cell.
int locality_order(bte_ctx* ctx) {
    ctx->ordered = true;
    ctx->cell_of.clear();
    const int N = ctx->N, F = ctx->F, dim = ctx->dim;
    if (dim < 2 || N < 4096) return BTE_OK;
    std::vector<std::vector<int>> coord;
    for (int a = 0; a < dim; a++) {
        int best = 0;
        for (int d = 1; d < ctx->D; d++)
            if (ctx->h_dir[3 * d + a] > ctx->h_dir[3 * best + a]) best = d;
        coord.push_back(upwind_levels(ctx, ctx->h_dir.data() + 3 * best));
    }
    auto spread = [](unsigned v, int dim) {  // bit i of v to bit dim*i
        unsigned long long r = 0;
        for (int i = 0; i < 21; i++) r |= (unsigned long long)((v >> i) & 1u) << (dim * i);
        return r;
    };
    std::vector<unsigned long long> key(N);
    for (int c = 0; c < N; c++) {
        unsigned long long k = 0;
        for (int a = 0; a < dim; a++) k |= spread((unsigned)coord[a][c], dim) << a;
        key[c] = k;
    }
    std::vector<int> cell_of(N), slot(N);
    for (int i = 0; i < N; i++) cell_of[i] = i;
    std::stable_sort(cell_of.begin(), cell_of.end(), [&](int x, int y) { return key[x] < key[y]; });
    for (int i = 0; i < N; i++) slot[cell_of[i]] = i;
    std::vector<double> vol(N), area((size_t)N * F), normal((size_t)N * F * 3);
    std::vector<int> nbr((size_t)N * F);
    for (int i = 0; i < N; i++) {
        const int c = cell_of[i];
        vol[i] = ctx->h_vol[c];
        for (int f = 0; f < F; f++) {
            const size_t ki = (size_t)i * F + f, kc = (size_t)c * F + f;
            area[ki] = ctx->h_area[kc];
            for (int a = 0; a < 3; a++) normal[3 * ki + a] = ctx->h_normal[3 * kc + a];
            const int nb = ctx->h_nbr[kc];
            nbr[ki] = nb >= 0 ? slot[nb] : nb;
        }
    }
    ctx->h_vol.swap(vol);
    ctx->h_area.swap(area);
    ctx->h_normal.swap(normal);
    ctx->h_nbr.swap(nbr);
    ctx->cell_of.swap(cell_of);
    HIPCHK(upload(ctx->vol, ctx->h_vol, ctx->stream));
    HIPCHK(upload(ctx->area, ctx->h_area, ctx->stream));
    HIPCHK(upload(ctx->normal, ctx->h_normal, ctx->stream));
    HIPCHK(upload(ctx->nbr, ctx->h_nbr, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->geo_tag = ++g_geo_tags;  // the network's graph of this mesh must be rebuilt
    ctx->learn_stat_ok = false;
    ctx->aa_cells_ok = false;
    return BTE_OK;
}

// Caller's index of internal cell i.
inline int ext_cell(const bte_ctx* ctx, int i) { return ctx->cell_of.empty() ? i : ctx->cell_of[i]; }

// Copies between the caller's cell order and the internal one: rows of `width` values per cell,
// `blocks` consecutive groups of N rows (e.g. [B][D][N] intensities: width 1, blocks B*D).
template <class T>
void to_internal(const bte_ctx* ctx, const T* ext, T* in, int width = 1, size_t blocks = 1) {
    const size_t N = ctx->N;
    for (size_t g = 0; g < blocks; g++)
        for (size_t i = 0; i < N; i++)
            for (int k = 0; k < width; k++)
                in[(g * N + i) * width + k] = ext[(g * N + ext_cell(ctx, (int)i)) * width + k];
}
template <class T>
void to_external(const bte_ctx* ctx, const T* in, T* ext, int width = 1, size_t blocks = 1) {
    const size_t N = ctx->N;
    for (size_t g = 0; g < blocks; g++)
        for (size_t i = 0; i < N; i++)
            for (int k = 0; k < width; k++)
                ext[(g * N + ext_cell(ctx, (int)i)) * width + k] = in[(g * N + i) * width + k];
}

// One-time preparation: cell order, boundary tables, assembly, sweep schedule, FP32 copies.
int prepare(bte_ctx* ctx, bool timing) {
    if (!ctx->have_geo || !ctx->have_quad || !ctx->have_bands || !ctx->have_bcs)
        SETERR(BTE_ERR_STATE, "bte_solve: geometry, quadrature, bands and boundary conditions must be set first");
    if (ctx->prepared) return BTE_OK;
    if (!ctx->ordered)
        if (int rc = locality_order(ctx)) return rc;
    const int N = ctx->N, F = ctx->F, D = ctx->D;
    // boundary faces + validation of the BC types they reference (StaticBTESolver.cpp:173-257)
    std::vector<int> bfc, bff, bfid((size_t)N * F, -1);
    ctx->has_diffuse = false;
    for (int c = 0; c < N; c++)
        for (int f = 0; f < F; f++) {
            const int nb = ctx->h_nbr[(size_t)c * F + f];
            if (nb >= 0) continue;
            bfid[(size_t)c * F + f] = (int)bfc.size();
            bfc.push_back(c);
            bff.push_back(f);
            for (size_t k = 0; k < ctx->h_bci.size(); k++) {
                if (ctx->h_bci[k] != nb) continue;
                const int t = ctx->h_bct[k];
                if (t == 2) ctx->has_diffuse = true;
                if (t == 1 || t == 2) continue;
                if (t == 31 || t == 32 || t == 33) {
                    for (int d = 0; d < D; d++) {
                        const int m = mirror_of(ctx->dim, ctx->num_phi, t, d);
                        if (m < 0 || m >= D)
                            SETERR(BTE_ERR_BC, "boundary condition type " + std::to_string(t) +
                                                   " not supported for this mesh dimension / quadrature");
                    }
                    continue;
                }
                SETERR(BTE_ERR_BC, "Boundary Condition Type " + std::to_string(t) + " not supported.");
            }
        }
    ctx->n_bf = (int)bfc.size();
    std::vector<int> mir(3 * (size_t)D, -1);
    for (int t = 0; t < 3; t++)
        for (int d = 0; d < D; d++) mir[(size_t)t * D + d] = mirror_of(ctx->dim, ctx->num_phi, 31 + t, d);
    HIPCHK(upload(ctx->bf_cell, bfc, ctx->stream));
    HIPCHK(upload(ctx->bf_face, bff, ctx->stream));
    HIPCHK(upload(ctx->bfid, bfid, ctx->stream));
    HIPCHK(upload(ctx->mirror, mir, ctx->stream));
    // effective angular weight of the volume term: omega_d in 2D/3D, 1 in 1D (StaticBTESolver.cpp:143-147)
    std::vector<double> weff(ctx->h_omega);
    if (ctx->dim == 1) std::fill(weff.begin(), weff.end(), 1.0);
    HIPCHK(upload(ctx->w_eff, weff, ctx->stream));

    // assembly
    const int W = std::max(1, F - 1);
    ctx->W = W;
    {
        Timer tm(ctx, KA_ASM, timing);
        HIPCHK(ctx->dg.ensure(sizeof(double) * (size_t)D * N));
        HIPCHK(ctx->ell_col.ensure(sizeof(int) * (size_t)D * W * N));
        HIPCHK(ctx->ell_val.ensure(sizeof(double) * (size_t)D * W * N));
        HIPCHK(ctx->overflow.ensure(sizeof(int)));
        HIPCHK(hipMemsetAsync(ctx->overflow.p, 0, sizeof(int), ctx->stream));
        hipLaunchKernelGGL(k_assemble, dim3(grid_for((long long)D * N, 256)), dim3(256), 0, ctx->stream, geo_of(ctx),
                           quad_of(ctx), W, ctx->dg.as<double>(), ctx->ell_col.as<int>(), ctx->ell_val.as<double>(),
                           ctx->overflow.as<int>());
        HIPCHK(hipGetLastError());
    }
    int ovf = 0;
    HIPCHK(hipMemcpyAsync(&ovf, ctx->overflow.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (ovf) SETERR(BTE_ERR_MESH, "mesh has cells with more incoming interior faces than F-1 (degenerate cells)");

    // sweep schedule: levels per direction on the host (threads over directions)
    {
        Timer tm(ctx, KA_SCHED, timing);
        std::vector<int> hcol((size_t)D * W * N);
        HIPCHK(hipMemcpyAsync(hcol.data(), ctx->ell_col.p, hcol.size() * sizeof(int), hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        std::vector<std::vector<int>> order(D), lptr(D), scol(D);
        std::vector<int> lag(D, 0);
        unsigned nth = std::max(1u, std::min<unsigned>({std::thread::hardware_concurrency(), 16u, (unsigned)D}));
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < nth; t++)
            pool.emplace_back([&, t]() {
                for (int d = (int)t; d < D; d += (int)nth)
                    level_schedule(N, W, hcol.data() + (size_t)d * W * N, order[d], lptr[d], scol[d], lag[d]);
            });
        for (auto& th : pool) th.join();
        std::vector<int> off(D), cnt(D), lp_all, perm_all((size_t)D * N), scol_all((size_t)D * W * N);
        std::vector<int> spos_all((size_t)D * W * N, -1);
        ctx->max_levels = 0;
        ctx->lagged = 0;
        for (int& r : ctx->ring_span) r = 0;
        ctx->max_back = 0;
        for (int d = 0; d < D; d++) {
            off[d] = (int)lp_all.size();
            cnt[d] = (int)lptr[d].size() - 1;
            ctx->max_levels = std::max(ctx->max_levels, cnt[d]);
            ctx->lagged += lag[d];
            lp_all.insert(lp_all.end(), lptr[d].begin(), lptr[d].end());
            std::copy(order[d].begin(), order[d].end(), perm_all.begin() + (size_t)d * N);
            std::copy(scol[d].begin(), scol[d].end(), scol_all.begin() + (size_t)d * W * N);
            // level-order positions of the upwind neighbours
            std::vector<int> inv(N);
            for (int i = 0; i < N; i++) inv[order[d][i]] = i;
            for (int l = 0; l + 1 < (int)lptr[d].size(); l++)
                for (int i = lptr[d][l]; i < lptr[d][l + 1]; i++)
                    for (int k = 0; k < W; k++) {
                        const int j = scol[d][(size_t)k * N + i];
                        if (j < 0) continue;
                        spos_all[((size_t)d * W + k) * N + i] = inv[j];
                        ctx->max_back = std::max(ctx->max_back, lptr[d][l + 1] - inv[j]);
                    }
            // positions spanned by K consecutive levels (the reach the ring needs when the sweep makes
            // HBM writes visible every K levels), K = 64, 32, ..., 2, 1
            for (int q = 0; q < 7; q++) {
                const int K = 64 >> q;
                for (int l = 0; l + 1 < (int)lptr[d].size(); l++)
                    ctx->ring_span[q] = std::max(ctx->ring_span[q], lptr[d][l + 1] - lptr[d][std::max(0, l - K + 1)]);
            }
        }
        // Level flags of the pipelined sweeps with a ring of ring_size(ctx) values: an upwind value
        // older than the ring is read from HBM (flag 1 on the reading level), so its write must
        // complete before a full work-group synchronisation that precedes the read, i.e. at the end
        // of some level in [level written, level read - 1] (flag 2).  Greedy interval stabbing
        // (intervals by their last level) flags the fewest level ends that cover every such read.
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) == hipSuccess && lds > 0)
            ctx->max_lds = lds;
        ctx->ring_r = ring_size(ctx);
        ctx->ring_r2 = ring_wide(ctx);
        std::vector<unsigned char> sync_all(lp_all.size(), 0);
        std::vector<int> fb_dir(D, 0);
        // the flags of a ring of R values into bits sel (reads) and sel + 1 (synchronise after), and
        // bit sel / 2 of fb_dir
        auto flag_levels = [&](int R, int sel) {
            std::vector<int> lv_of(N);
            std::vector<std::pair<int, int>> iv;  // (last, first) level end that may synchronise
            for (int d = 0; d < D; d++) {
                const std::vector<int>& P = lptr[d];
                const int L = (int)P.size() - 1;
                for (int l = 0; l < L; l++)
                    for (int i = P[l]; i < P[l + 1]; i++) lv_of[i] = l;
                iv.clear();
                for (int l = 0; l < L; l++)
                    for (int i = P[l]; i < P[l + 1]; i++)
                        for (int k = 0; k < W; k++) {
                            const int p = spos_all[((size_t)d * W + k) * N + i];
                            if (p >= 0 && p < P[l + 1] - R) {
                                iv.push_back({l - 1, lv_of[p]});
                                sync_all[off[d] + l] |= (unsigned char)(1 << sel);
                            }
                        }
                if (iv.empty()) continue;
                fb_dir[d] |= 1 << (sel / 2);
                std::sort(iv.begin(), iv.end());
                int last = -1;
                for (const auto& e : iv)
                    if (last < e.second) {
                        last = e.first;
                        sync_all[off[d] + last] |= (unsigned char)(2 << sel);
                    }
            }
        };
        if (ctx->ring_r > 0) flag_levels(ctx->ring_r, 0);
        if (ctx->ring_r2 > 0) flag_levels(ctx->ring_r2, 2);
        HIPCHK(upload(ctx->lvl_sync, sync_all, ctx->stream));
        HIPCHK(upload(ctx->dir_fb, fb_dir, ctx->stream));
        ctx->h_lvl_cnt = cnt;
        HIPCHK(upload(ctx->lvl_off, off, ctx->stream));
        HIPCHK(upload(ctx->lvl_cnt, cnt, ctx->stream));
        HIPCHK(upload(ctx->lvl_ptr, lp_all, ctx->stream));
        HIPCHK(upload(ctx->perm, perm_all, ctx->stream));
        HIPCHK(upload(ctx->swp_col, scol_all, ctx->stream));
        HIPCHK(upload(ctx->swp_pos, spos_all, ctx->stream));
        HIPCHK(ctx->swp_dg.ensure(sizeof(double) * (size_t)D * N));
        HIPCHK(ctx->swp_val.ensure(sizeof(double) * (size_t)D * W * N));
        hipLaunchKernelGGL(k_gather_sweep<double>, dim3(grid_for((long long)D * N, 256)), dim3(256), 0, ctx->stream, N,
                           D, W, ctx->perm.as<int>(), ctx->dg.as<double>(), ctx->ell_val.as<double>(),
                           ctx->swp_dg.as<double>(), ctx->swp_val.as<double>());
        HIPCHK(ctx->swp_vw.ensure(sizeof(double) * (size_t)D * N));
        HIPCHK(ctx->swp_bnd.ensure((size_t)D * N));
        hipLaunchKernelGGL(k_sweep_aux, dim3(grid_for((long long)D * N, 256)), dim3(256), 0, ctx->stream, geo_of(ctx),
                           quad_of(ctx), ctx->perm.as<int>(), ctx->w_eff.as<double>(), ctx->swp_vw.as<double>(),
                           ctx->swp_bnd.as<unsigned char>());
        HIPCHK(hipGetLastError());
        HIPCHK(ctx->swp_pki.ensure(sizeof(int) * 4 * (size_t)D * N));
        HIPCHK(ctx->swp_pkd.ensure(sizeof(double) * 4 * (size_t)D * N));
        hipLaunchKernelGGL(k_pack_sweep, dim3(grid_for((long long)D * N, 256)), dim3(256), 0, ctx->stream, N, D, W,
                           ctx->perm.as<int>(), ctx->swp_pos.as<int>(), ctx->swp_val.as<double>(),
                           ctx->swp_dg.as<double>(), ctx->swp_vw.as<double>(), ctx->swp_pki.as<int>(),
                           ctx->swp_pkd.as<double>());
        HIPCHK(hipGetLastError());
        // compact list of the level-order entries with an incoming boundary face
        std::vector<unsigned char> hb((size_t)D * N);
        HIPCHK(hipMemcpyAsync(hb.data(), ctx->swp_bnd.p, hb.size(), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        std::vector<int> blist;
        for (size_t t = 0; t < hb.size(); t++)
            if (hb[t]) blist.push_back((int)t);
        ctx->n_bnd_list = (int)blist.size();
        HIPCHK(upload(ctx->bnd_list, blist, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    ctx->have_sweep = true;
    ctx->have_fp32 = false;
    ctx->prepared = true;
    drop_graphs(ctx);
    return BTE_OK;
}

int ensure_fp32(bte_ctx* ctx) {
    if (ctx->have_fp32) return BTE_OK;
    const int N = ctx->N, D = ctx->D, W = ctx->W;
    HIPCHK(ctx->dg32.ensure(sizeof(float) * (size_t)D * N));
    HIPCHK(ctx->ell_val32.ensure(sizeof(float) * (size_t)D * W * N));
    HIPCHK(ctx->swp_dg32.ensure(sizeof(float) * (size_t)D * N));
    HIPCHK(ctx->swp_val32.ensure(sizeof(float) * (size_t)D * W * N));
    hipLaunchKernelGGL(k_convert<float>, dim3(grid_for((long long)D * N, 256)), dim3(256), 0, ctx->stream,
                       (long long)D * N, ctx->dg.as<double>(), ctx->dg32.as<float>());
    hipLaunchKernelGGL(k_convert<float>, dim3(grid_for((long long)D * W * N, 256)), dim3(256), 0, ctx->stream,
                       (long long)D * W * N, ctx->ell_val.as<double>(), ctx->ell_val32.as<float>());
    hipLaunchKernelGGL(k_convert<float>, dim3(grid_for((long long)D * N, 256)), dim3(256), 0, ctx->stream,
                       (long long)D * N, ctx->swp_dg.as<double>(), ctx->swp_dg32.as<float>());
    hipLaunchKernelGGL(k_convert<float>, dim3(grid_for((long long)D * W * N, 256)), dim3(256), 0, ctx->stream,
                       (long long)D * W * N, ctx->swp_val.as<double>(), ctx->swp_val32.as<float>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->have_fp32 = true;
    return BTE_OK;
}

// Device buffers of an iteration, sized from the current inputs: intensities and Krylov vectors
// (B*D*N), per-system counters (B*D), per-(band, cell) and per-cell fields, the heat flux per (BC,
// band), the diffuse re-emission per (band, boundary face) and the margin reduction.  Each is
// checked every time (a no-op when it is large enough): the setters change B, D, N, the BC count
// and the boundary faces independently.  A buffer that had to grow moves, so graphs captured with
// the old address are dropped.  During bte_solve_noalloc the reduction buffers are the device's
// scratch slices and the context's own are not needed.
int ensure_work(bte_ctx* ctx, int precision) {
    const size_t SN = (size_t)ctx->B * ctx->D * ctx->N, BN = (size_t)ctx->B * ctx->N;
    const int S = ctx->B * ctx->D;
    // intensities: room for either layout (the interleaved one pads the bands to a multiple of 4);
    // Krylov vectors are FP64-sized in both precisions: the mixed form re-solves in FP64 a system
    // whose FP32 inner solve stalls; the right-hand-side buffer also holds the level-order rows of
    // stride lo_stride(N)
    const size_t EE = std::max(SN, (size_t)ctx->D * ctx->N * ee_padded_bands(ctx->B));
    set_layout(ctx, false);
    std::vector<std::pair<DBuf*, size_t>> want = {
        {&ctx->eeA, 8 * EE}, {&ctx->eeB, 8 * EE}, {&ctx->wb, 8 * (size_t)S * lo_stride(ctx->N)},
        {&ctx->wr, 8 * SN}, {&ctx->wrh, 8 * SN},
        {&ctx->wp, 8 * SN}, {&ctx->wv, 8 * SN}, {&ctx->wph, 8 * SN}, {&ctx->wsh, 8 * SN}, {&ctx->wt, 8 * SN},
        {&ctx->iters, sizeof(int) * S}, {&ctx->fails, sizeof(int) * S}, {&ctx->dens, 8 * BN},
        {&ctx->T, 8 * (size_t)ctx->N}, {&ctx->Tband, 8 * BN}, {&ctx->tsrc, 8 * (size_t)ctx->N},
        {&ctx->tlo, 8 * (size_t)ctx->D * lo_stride(ctx->N)},
        {&ctx->ediff, 8 * std::max<size_t>(1, (size_t)ctx->B * ctx->n_bf)}};
    if (precision == BTE_PREC_MIXED) {
        want.push_back({&ctx->wr64, 8 * SN});
        want.push_back({&ctx->wy, 4 * SN});
        want.push_back({&ctx->wf, 4 * SN});
    }
    if (!ctx->use_scratch) {
        want.push_back({&ctx->partial, 8 * (size_t)grid_for((long long)BN, 256)});
        want.push_back({&ctx->margin, 8 * 2});
        want.push_back({&ctx->bflux, 8 * std::max<size_t>(1, ctx->h_bci.size() * ctx->B)});
    }
    bool moved = false;
    for (auto& w : want) {
        const void* before = w.first->p;
        HIPCHK(w.first->ensure(w.second));
        moved |= before != w.first->p;
    }
    if (moved) drop_graphs(ctx);
    return BTE_OK;
}

Ops ops_of(bte_ctx* ctx) {
    Ops o{};
    o.N = ctx->N;
    o.D = ctx->D;
    o.W = ctx->W;
    o.Np = lo_stride(ctx->N);
    o.vol = ctx->vol.as<double>();
    o.omega = ctx->w_eff.as<double>();
    o.vt = ctx->vt.as<double>();
    o.band_C = ctx->bC.as<double>();
    o.solid_angle = ctx->solid_angle;
    o.T_ref = ctx->T_ref;
    o.dg = ctx->dg.as<double>();
    o.ell_col = ctx->ell_col.as<int>();
    o.ell_val = ctx->ell_val.as<double>();
    o.ell_val32 = ctx->ell_val32.as<float>();
    o.dg32 = ctx->dg32.as<float>();
    o.lvl_off = ctx->lvl_off.as<int>();
    o.lvl_cnt = ctx->lvl_cnt.as<int>();
    o.lvl_ptr = ctx->lvl_ptr.as<int>();
    o.perm = ctx->perm.as<int>();
    o.swp_col = ctx->swp_col.as<int>();
    o.swp_val = ctx->swp_val.as<double>();
    o.swp_dg = ctx->swp_dg.as<double>();
    o.swp_val32 = ctx->swp_val32.as<float>();
    o.swp_dg32 = ctx->swp_dg32.as<float>();
    o.swp_vw = ctx->swp_vw.as<double>();
    o.swp_bnd = ctx->swp_bnd.as<unsigned char>();
    o.swp_pos = ctx->swp_pos.as<int>();
    o.lvl_sync = ctx->lvl_sync.as<unsigned char>();
    o.dir_fb = ctx->dir_fb.as<int>();
    o.ring_r = ctx->ring_r;
    o.ring_r2 = ctx->ring_r2;
    o.ring_sel = 0;
    o.swp_pki = ctx->swp_pki.as<int>();
    o.swp_pkd = ctx->swp_pkd.as<double>();
    o.ring_sync = 1;
    (void)ring_size(ctx, &o.ring_sync);
    o.ee = ctx->eei;
    return o;
}

Work work_of(bte_ctx* ctx, double* ee_curr) {
    Work w{};
    w.x = ee_curr;
    w.b = ctx->wb.as<double>();
    w.tlo = ctx->tlo.as<double>();
    w.r64 = ctx->wr64.as<double>();
    w.r = ctx->wr.p;
    w.rh = ctx->wrh.p;
    w.p = ctx->wp.p;
    w.v = ctx->wv.p;
    w.ph = ctx->wph.p;
    w.sh = ctx->wsh.p;
    w.t = ctx->wt.p;
    w.y = ctx->wy.p;
    w.f = ctx->wf.p;
    w.iters = ctx->iters.as<int>();
    w.fails = ctx->fails.as<int>();
    return w;
}

// Launch shape of the inner-solve kernel: work-group size and, for the direct sweep, its variant
// (1 basic, 2 level-ordered operands, 3 level-ordered operands + intensity vector in LDS,
// 4 level-ordered operands + LDS ring of recent intensities + pipelined operand loads, 5 as 4 with
// one wave per system, 6 / 7 as 4 with 2 / 4 bands per work-group, 8 as 4 with the wide ring).
struct Shape {
    int block = 256;
    int variant = 2;  // direct form: sweep variant; Krylov forms: bands per work-group (1, 2, 4)
};

// Bands per work-group a Krylov form uses when it is not autotuned: opts.band_group if it is a
// supported group of the Jacobi path, else one system per work-group.
int krylov_group(const bte_opts* o) {
    if (o->precond != BTE_PC_JACOBI || o->precision != BTE_PREC_FP64) return 1;
    return (o->band_group == 2 || o->band_group == 4) ? o->band_group : 1;
}

// Whether the band-batched kernels run G bands per work-group of `block` threads: the FP64 Jacobi
// path only (for the FP32 inner solves of the mixed form one system per work-group measured
// faster), G in {2, 4}, work-groups of at most 512 (wider ones would spill), and G bands of D*N
// FP64 elements below 4 GiB (the kernels address a group's vectors through 32-bit buffer offsets).
bool group_ok(const bte_ctx* ctx, const bte_opts* o, int G, int block) {
    return o->precond == BTE_PC_JACOBI && o->precision == BTE_PREC_FP64 && (G == 2 || G == 4) && block <= 512 &&
           (double)G * ctx->D * ctx->N * sizeof(double) < 4294967295.0;
}

bool direct_path(const bte_ctx* ctx, const bte_opts* o) {
    return o->precision == BTE_PREC_FP64 && o->precond == BTE_PC_SWEEP && o->direct_sweep && ctx->lagged == 0;
}

bool lds_fits(const bte_ctx* ctx) { return (size_t)ctx->N * sizeof(double) + 1024 <= (size_t)ctx->max_lds; }

// LDS ring of the pipelined sweep (variant 4): 8192 entries (64 KB; the fewer values are older than
// the ring, the fewer reads go to HBM), or the smallest power of two >= 1024 that still fits the LDS
// with the level pointers; it must cover the positions of K consecutive levels, K the largest of
// {64, 32, ..., 2, 1} that it covers (the fewer full synchronisations, the longer the prefetches
// stay in flight); *sync_every = K.  0 when even one level is wider than the ring.
// With G bands per work-group (variants 6, 7) each band has a ring of 8192 / G entries (the same
// LDS in total), at least 1024.
int ring_size(const bte_ctx* ctx, int* sync_every, int G) {
    if (ctx->lagged != 0) return 0;
    int r = 8192 / G;
    while (r > 1024 && ring_lds_bytes(G * r, ctx->max_levels) > (size_t)ctx->max_lds) r >>= 1;
    if (r < 1024 || ring_lds_bytes(G * r, ctx->max_levels) > (size_t)ctx->max_lds) return 0;
    for (int q = 0; q < 7; q++)
        if (ctx->ring_span[q] > 0 && ctx->ring_span[q] <= r) {
            if (sync_every) *sync_every = 64 >> q;
            return r;
        }
    return 0;
}

// The wide ring of the pipelined sweep (variant 8): twice the largest ring, 16,384 values (128 KB of
// LDS, so one work-group per compute unit), where it fits; 0 otherwise.  Fewer upwind values are
// then older than the ring and read back from HBM, which pays where there are many more systems
// than compute units.
int ring_wide(const bte_ctx* ctx) {
    const int r = ring_size(ctx);
    if (r != 8192 || ring_lds_bytes(2 * r, ctx->max_levels) > (size_t)ctx->max_lds) return 0;
    return 2 * r;
}

// Right-hand sides of all systems in level order (interior cells, then the boundary-cell list);
// with_diffuse: the diffuse re-emission of ee_prev first, in the interior pass's launch.
void enqueue_rhs_levels(bte_ctx* ctx, hipStream_t st, const double* ee_prev, const double* dens, double* out,
                        bool with_diffuse = false) {
    const int S = ctx->B * ctx->D;
    const int n_rhs = grid_for((long long)ctx->D * lo_stride(ctx->N), 256);
    // the interior cells' densities from their source temperatures while those hold
    const double* ts = (ctx->tsrc_ok && dens == ctx->dens.as<double>()) ? ctx->tsrc.as<double>() : nullptr;
    if (with_diffuse) {
        const int n_diff = diffuse_grid(ctx);
        hipLaunchKernelGGL(k_diffuse_rhs_levels, dim3((unsigned)(n_diff + n_rhs)), dim3(256), 0, st, geo_of(ctx),
                           quad_of(ctx), bnd_of(ctx), ops_of(ctx), ctx->B, ee_prev, ctx->ediff.as<double>(), dens,
                           ts, out, n_diff);
    } else {
        hipLaunchKernelGGL(k_rhs_levels, dim3(n_rhs), dim3(256), 0, st, ops_of(ctx), ctx->B, dens, ts, out);
    }
    if (ctx->n_bnd_list > 0)
        hipLaunchKernelGGL(k_rhs_levels_bnd, dim3(rhs_bnd_grid(ctx, kBndEntries)), dim3(kBndEntries), 0, st,
                           geo_of(ctx),
                           quad_of(ctx), bnd_of(ctx), bands_of(ctx), ops_of(ctx), S, ctx->bnd_list.as<int>(),
                           ctx->n_bnd_list, ee_prev, dens, out);
}

// For the sweeps that form the interior right-hand sides themselves: the level-order source
// temperatures (with the diffuse re-emission in the same launch when with_diffuse), then the
// right-hand sides of the cells with an incoming boundary face into out, as enqueue_rhs_levels.
void enqueue_tsrc_levels(bte_ctx* ctx, hipStream_t st, const double* ee_prev, const double* dens, double* out,
                         bool with_diffuse = false) {
    const int S = ctx->B * ctx->D;
    const int n_lo = grid_for((long long)ctx->D * lo_stride(ctx->N), 256);
    if (with_diffuse) {
        const int n_diff = diffuse_grid(ctx);
        hipLaunchKernelGGL(k_diffuse_tsrc_levels, dim3((unsigned)(n_diff + n_lo)), dim3(256), 0, st, geo_of(ctx),
                           quad_of(ctx), bnd_of(ctx), ops_of(ctx), ctx->B, ee_prev, ctx->ediff.as<double>(),
                           ctx->tsrc.as<double>(), ctx->tlo.as<double>(), n_diff);
    } else {
        hipLaunchKernelGGL(k_tsrc_levels, dim3(n_lo), dim3(256), 0, st, ops_of(ctx), ctx->tsrc.as<double>(),
                           ctx->tlo.as<double>());
    }
    if (ctx->n_bnd_list > 0)
        hipLaunchKernelGGL(k_rhs_levels_bnd, dim3(rhs_bnd_grid(ctx, kBndEntries)), dim3(kBndEntries), 0, st, geo_of(ctx), quad_of(ctx),
                           bnd_of(ctx), bands_of(ctx), ops_of(ctx), S, ctx->bnd_list.as<int>(), ctx->n_bnd_list,
                           ee_prev, dens, out);
}

// marks (null, or three events): recorded before the right-hand sides, between them and the inner
// solve, and after it
inline void mark(hipEvent_t* marks, int k, hipStream_t st) {
    if (marks) (void)hipEventRecord(marks[k], st);
}

template <int BLOCK>
void launch_solve_b(bte_ctx* ctx, const bte_opts* o, int variant, int S, const double* ee_prev, double* ee_curr,
                    const double* dens, bool diffuse_first, hipEvent_t* marks) {
    SolveCfg cfg{S, o->inner_maxit, o->inner_tol, std::max(1, o->refine_steps), o->precond == BTE_PC_JACOBI ? 1 : 0,
                 o->inner_tol32 > 0 ? o->inner_tol32 : 1e-5};
    const Geo g = geo_of(ctx);
    const Quad q = quad_of(ctx);
    const Bnd bd = bnd_of(ctx);
    const Bands bn = bands_of(ctx);
    const Ops op = ops_of(ctx);
    const Work w = work_of(ctx, ee_curr);
    dim3 grid(S), blk(BLOCK);
    // diffuse_first: the diffuse re-emission of ee_prev is still due; the pipelined sweep's
    // right-hand-side pass takes it into its launch, every other form runs it first
    const bool v4 = direct_path(ctx, o) && ((variant == 4 && ring_size(ctx) > 0) || (variant == 8 && ctx->ring_r2 > 0));
    mark(marks, 0, ctx->stream);
    if (diffuse_first && !v4)
        hipLaunchKernelGGL(k_diffuse, dim3(diffuse_grid(ctx)), dim3(256), 0, ctx->stream,
                           g, q, bd, ctx->B, ee_prev, ctx->ediff.as<double>());
    if (direct_path(ctx, o)) {
        int sync_every = 1;
        const int ring = ring_size(ctx, &sync_every);
        if (variant == 5 && ring > 0) {
            ctx->variant_used = 5;
            enqueue_rhs_levels(ctx, ctx->stream, ee_prev, dens, w.b);
            mark(marks, 1, ctx->stream);
            const size_t lds = ring_lds_bytes(ring, ctx->max_levels);
            auto go = [&](auto kernel) {
                (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                hipLaunchKernelGGL(kernel, grid, dim3(64), lds, ctx->stream, op, w, S, ring - 1);
            };
            // without reads older than the ring the sweep needs no HBM read-back path
            const bool fb = ctx->max_back > ring;
            if (ctx->W == 1)
                fb ? go(k_sweep_wave<1, true>) : go(k_sweep_wave<1, false>);
            else if (ctx->W == 2)
                fb ? go(k_sweep_wave<2, true>) : go(k_sweep_wave<2, false>);
            else
                fb ? go(k_sweep_wave<3, true>) : go(k_sweep_wave<3, false>);
        } else if ((variant == 6 || variant == 7) && ctx->B >= (variant == 6 ? 2 : 4) &&
                   ring_size(ctx, nullptr, variant == 6 ? 2 : 4) > 0) {
            // G bands of a direction per work-group, each with its own ring of 8192 / G entries
            const int G = variant == 6 ? 2 : 4;
            int sync_g = 1;
            const int ring_g = ring_size(ctx, &sync_g, G);
            ctx->variant_used = variant;
            enqueue_rhs_levels(ctx, ctx->stream, ee_prev, dens, w.b);
            mark(marks, 1, ctx->stream);
            Ops opg = op;
            opg.ring_sync = sync_g;
            const size_t lds = ring_lds_bytes(G * ring_g, ctx->max_levels);
            const dim3 gg(ctx->D * ((ctx->B + G - 1) / G));
            auto go = [&](auto kernel) {
                (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                hipLaunchKernelGGL(kernel, gg, blk, lds, ctx->stream, opg, w, ctx->B, ring_g - 1);
            };
            const bool fb = ctx->max_back > ring_g;
            auto by_w = [&](auto gc) {
                constexpr int GG = decltype(gc)::value;
                if (ctx->W == 1)
                    fb ? go(k_sweep_ring_bands<BLOCK, 1, GG, true>) : go(k_sweep_ring_bands<BLOCK, 1, GG, false>);
                else if (ctx->W == 2)
                    fb ? go(k_sweep_ring_bands<BLOCK, 2, GG, true>) : go(k_sweep_ring_bands<BLOCK, 2, GG, false>);
                else
                    fb ? go(k_sweep_ring_bands<BLOCK, 3, GG, true>) : go(k_sweep_ring_bands<BLOCK, 3, GG, false>);
            };
            if (G == 2)
                by_w(Unroll<2>{});
            else
                by_w(Unroll<4>{});
        } else if ((variant == 4 && ring > 0) || (variant == 8 && ctx->ring_r2 > 0)) {
            // variant 8: the wide ring, with the level flags computed for it
            const bool wide = variant == 8;
            const int R = wide ? ctx->ring_r2 : ring;
            ctx->variant_used = variant;
            Ops opr = op;
            opr.ring_sel = wide ? 1 : 0;
            // right-hand sides in level order into w.b, then the pipelined sweep.  Variant 8, while
            // the source temperatures hold: only the boundary cells' ones, the sweep forming the
            // interior ones (a few more instructions per cell, which the issue-bound sweep of a
            // launch with few systems per compute unit, variant 4's case, does not recover)
            const bool ts = wide && ctx->tsrc_ok && dens == ctx->dens.as<double>();
            if (ts)
                enqueue_tsrc_levels(ctx, ctx->stream, ee_prev, dens, w.b, diffuse_first);
            else
                enqueue_rhs_levels(ctx, ctx->stream, ee_prev, dens, w.b, diffuse_first);
            mark(marks, 1, ctx->stream);
            const size_t lds = ring_lds_bytes(R, ctx->max_levels);
            auto go = [&](auto kernel) {
                (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                hipLaunchKernelGGL(kernel, grid, blk, lds, ctx->stream, opr, w, S, R - 1);
            };
            const bool fb = ctx->max_back > R;
            auto by_w = [&](auto tsc) {
                constexpr bool TS = decltype(tsc)::value;
                if (ctx->W == 1)
                    fb ? go(k_sweep_ring<BLOCK, 1, true, TS>) : go(k_sweep_ring<BLOCK, 1, false, TS>);
                else if (ctx->W == 2)
                    fb ? go(k_sweep_ring<BLOCK, 2, true, TS>) : go(k_sweep_ring<BLOCK, 2, false, TS>);
                else
                    fb ? go(k_sweep_ring<BLOCK, 3, true, TS>) : go(k_sweep_ring<BLOCK, 3, false, TS>);
            };
            if (ts)
                by_w(BoolC<true>{});
            else
                by_w(BoolC<false>{});
        } else if (variant == 3 && lds_fits(ctx)) {
            ctx->variant_used = 3;
            mark(marks, 1, ctx->stream);  // these sweeps form their right-hand sides themselves
            const size_t lds = (size_t)ctx->N * sizeof(double);
            (void)hipFuncSetAttribute((const void*)k_sweep_fast<BLOCK, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds);
            hipLaunchKernelGGL((k_sweep_fast<BLOCK, true>), grid, blk, lds, ctx->stream, g, q, bd, bn, op, w, S, ee_prev,
                               dens);
        } else if (variant == 1) {
            ctx->variant_used = 1;
            mark(marks, 1, ctx->stream);
            hipLaunchKernelGGL((k_sweep_direct<BLOCK>), grid, blk, 0, ctx->stream, g, q, bd, bn, op, w, S, ee_prev, dens);
        } else {
            ctx->variant_used = 2;
            mark(marks, 1, ctx->stream);
            hipLaunchKernelGGL((k_sweep_fast<BLOCK, false>), grid, blk, 0, ctx->stream, g, q, bd, bn, op, w, S, ee_prev,
                               dens);
        }
        mark(marks, 2, ctx->stream);
        return;
    }
    // Krylov forms: the right-hand sides of all systems first (keeps the face loop's registers out
    // of the solve kernel, whose occupancy sets its bandwidth)
    hipLaunchKernelGGL(k_rhs, dim3(grid_for((long long)S * ctx->N, 256)), dim3(256), 0, ctx->stream, g, q, bd, bn,
                       ee_prev, dens, w.b);
    mark(marks, 1, ctx->stream);
    const int G = group_ok(ctx, o, variant, BLOCK) ? variant : 1;
    if constexpr (BLOCK <= 512) if (G > 1) {
        // band-batched: one work-group per (direction, group of G bands)
        const dim3 gg(ctx->D * ((ctx->B + G - 1) / G));
        if (G == 2)
            hipLaunchKernelGGL((k_solve64_bands<BLOCK, 2>), gg, blk, 0, ctx->stream, op, w, cfg);
        else
            hipLaunchKernelGGL((k_solve64_bands<BLOCK, 4>), gg, blk, 0, ctx->stream, op, w, cfg);
        mark(marks, 2, ctx->stream);
        return;
    }
    if (o->precision == BTE_PREC_MIXED) {
        if (o->precond == BTE_PC_JACOBI)
            hipLaunchKernelGGL((k_solve_mixed<BLOCK, false>), grid, blk, 0, ctx->stream, op, w, cfg);
        else
            hipLaunchKernelGGL((k_solve_mixed<BLOCK, true>), grid, blk, 0, ctx->stream, op, w, cfg);
    } else {
        if (o->precond == BTE_PC_JACOBI)
            hipLaunchKernelGGL((k_solve64<BLOCK, false>), grid, blk, 0, ctx->stream, op, w, cfg);
        else
            hipLaunchKernelGGL((k_solve64<BLOCK, true>), grid, blk, 0, ctx->stream, op, w, cfg);
    }
    mark(marks, 2, ctx->stream);
}

void launch_solve(bte_ctx* ctx, const bte_opts* o, Shape sh, const double* ee_prev, double* ee_curr,
                  const double* dens, bool diffuse_first = false, hipEvent_t* marks = nullptr) {
    const int S = ctx->B * ctx->D;
    switch (sh.block) {
        case 64: launch_solve_b<64>(ctx, o, sh.variant, S, ee_prev, ee_curr, dens, diffuse_first, marks); break;
        case 128: launch_solve_b<128>(ctx, o, sh.variant, S, ee_prev, ee_curr, dens, diffuse_first, marks); break;
        case 512: launch_solve_b<512>(ctx, o, sh.variant, S, ee_prev, ee_curr, dens, diffuse_first, marks); break;
        case 1024: launch_solve_b<1024>(ctx, o, sh.variant, S, ee_prev, ee_curr, dens, diffuse_first, marks); break;
        default: launch_solve_b<256>(ctx, o, sh.variant, S, ee_prev, ee_curr, dens, diffuse_first, marks); break;
    }
}

// The kernels of one outer iteration (diffuse BC data, inner solves, T update, margin).
int enqueue_iteration(bte_ctx* ctx, const bte_opts* o, Shape sh, const double* ee_prev, double* ee_curr,
                      bool timing) {
    const Bands bn = bands_of(ctx);
    if (timing) {
        // the right-hand sides (with the diffuse re-emission of ee_prev, which only they read) and
        // the inner-solve kernel timed apart with events on the library stream
        hipEvent_t marks[3];
        for (auto& e : marks) HIPCHK(hipEventCreate(&e));
        launch_solve(ctx, o, sh, ee_prev, ee_curr, ctx->dens.as<double>(), ctx->has_diffuse, marks);
        HIPCHK(hipEventSynchronize(marks[2]));
        float t01 = 0, t12 = 0;
        HIPCHK(hipEventElapsedTime(&t01, marks[0], marks[1]));
        HIPCHK(hipEventElapsedTime(&t12, marks[1], marks[2]));
        for (auto& e : marks) (void)hipEventDestroy(e);
        ctx->ms[KA_RHS] += t01;
        ctx->calls[KA_RHS]++;
        ctx->ms[KA_SOLVE] += t12;
        ctx->calls[KA_SOLVE]++;
    } else {
        // the diffuse re-emission of ee_prev, which only the right-hand sides read, goes with the
        // solve launch (beside the pipelined sweep's right-hand-side pass, or just before)
        launch_solve(ctx, o, sh, ee_prev, ee_curr, ctx->dens.as<double>(), ctx->has_diffuse);
    }
    // interleaved intensities: two (cell, band) pairs per thread, so that a work-group reads each
    // direction's plane in runs of 4 KB (3D 200k: 1.55 -> 1.30 ms)
    const int pairs_per_thread = ctx->ee_inter ? 2 : 1;
    const int nblk = grid_for(((long long)ctx->B * ctx->N + pairs_per_thread - 1) / pairs_per_thread, 256);
    {
        Timer tm(ctx, KA_TEMP, timing);
        if (ctx->ee_inter)
            hipLaunchKernelGGL((k_tband<256, true, 2>), dim3(nblk), dim3(256), 0, ctx->stream, ctx->N, ctx->D, bn,
                               ctx->eei, ctx->omega.as<double>(), ee_curr, ee_prev, ctx->Tband.as<double>(),
                               partial_of(ctx));
        else
            hipLaunchKernelGGL((k_tband<256, false>), dim3(nblk), dim3(256), 0, ctx->stream, ctx->N, ctx->D, bn,
                               ctx->eei, ctx->omega.as<double>(), ee_curr, ee_prev, ctx->Tband.as<double>(),
                               partial_of(ctx));
        // cell temperatures and densities, and the margin, in one launch (both read the band step)
        const int n_tcell = grid_for(ctx->N, 256);
        hipLaunchKernelGGL(k_tcell_margin, dim3(n_tcell + 1), dim3(256), 0, ctx->stream, ctx->N, bn,
                           ctx->Tband.as<double>(), ctx->T.as<double>(), ctx->dens.as<double>(),
                           ctx->tsrc.as<double>(), n_tcell, nblk,
                           partial_of(ctx), 1.0 / ctx->B / ctx->N, margin_of(ctx), ctx->d_sync);
    }
    HIPCHK(hipGetLastError());
    return BTE_OK;
}

// Autotuner: times the inner-solve kernel of one outer iteration (from a zero state) for each
// candidate work-group size and sweep variant, keeps the fastest for this solver form.
int autotune(bte_ctx* ctx, const bte_opts* o, Shape& best) {
    const bool direct = direct_path(ctx, o);
    std::vector<Shape> cand;
    for (int blk : {64, 128, 256, 512, 1024}) {
        if (direct) {
            if (blk == 1024) continue;
            cand.push_back({blk, 2});
            if (lds_fits(ctx)) cand.push_back({blk, 3});
            if (ring_size(ctx) > 0) cand.push_back({blk, 4});
            if (blk >= 128 && ctx->ring_r2 > 0) cand.push_back({blk, 8});
            if (blk == 64 && ring_size(ctx) > 0) cand.push_back({64, 5});  // one wave per system
            if (blk >= 128 && ctx->B >= 2 && ring_size(ctx, nullptr, 2) > 0) cand.push_back({blk, 6});
            if (blk >= 128 && ctx->B >= 4 && ring_size(ctx, nullptr, 4) > 0) cand.push_back({blk, 7});
        } else if (blk >= 128) {
            for (int G : {1, 2, 4}) {
                const bool allowed = (o->band_group == 0) ? (G == 1 || (G <= ctx->B && group_ok(ctx, o, G, blk)))
                                                          : G == krylov_group(o);
                if (allowed) cand.push_back({blk, G});
            }
        }
    }
    // Krylov forms are timed on a fixed number of iterations per system: the rate per iteration is
    // what the shape changes, and a full solve from the zero state can take seconds per candidate.
    bte_opts ot = *o;
    if (!direct) ot.inner_maxit = std::min(o->inner_maxit, 12);
    o = &ot;
    const size_t SN = (size_t)ctx->B * ctx->D * ctx->N;
    HIPCHK(hipMemsetAsync(ctx->eeA.p, 0, sizeof(double) * SN, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->dens.p, 0, sizeof(double) * (size_t)ctx->B * ctx->N, ctx->stream));
    hipLaunchKernelGGL(k_fill, dim3(grid_for(ctx->N, 256)), dim3(256), 0, ctx->stream, ctx->N, ctx->T_ref,
                       ctx->tsrc.as<double>());  // the zero density's source temperature
    ctx->tsrc_ok = true;
    if (ctx->has_diffuse)
        hipLaunchKernelGGL(k_diffuse, dim3(diffuse_grid(ctx)), dim3(256), 0, ctx->stream,
                           geo_of(ctx), quad_of(ctx), bnd_of(ctx), ctx->B, ctx->eeA.as<double>(), ctx->ediff.as<double>());
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    float best_ms = 1e30f;
    for (const Shape& sh : cand) {
        float tot = 0;
        for (int rep = 0; rep < 3; rep++) {
            HIPCHK(hipEventRecord(e0, ctx->stream));
            launch_solve(ctx, o, sh, ctx->eeA.as<double>(), ctx->eeB.as<double>(), ctx->dens.as<double>());
            HIPCHK(hipEventRecord(e1, ctx->stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            if (rep > 0) tot += ms;  // first launch warms caches and code
        }
        if (tot < best_ms) {
            best_ms = tot;
            best = sh;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(hipGetLastError());
    return BTE_OK;
}

// allow_tune: the autotuner overwrites the intensity buffers and the density, so it only runs where
// that state is initialised afterwards (bte_solve); elsewhere an untuned form uses {256, 2}.
int pick_shape(bte_ctx* ctx, const bte_opts* o, Shape& sh, bool allow_tune) {
    const bool direct = direct_path(ctx, o);
    const int gkey = direct ? 0 : (o->band_group == 1 ? 1 : o->band_group == 2 ? 2 : o->band_group == 4 ? 3 : 0);
    int& tb = ctx->tuned_block[o->precision == BTE_PREC_MIXED][o->precond == BTE_PC_JACOBI][direct][gkey];
    int& tv = ctx->tuned_variant[o->precision == BTE_PREC_MIXED][o->precond == BTE_PC_JACOBI][direct][gkey];
    const bool fixed_block = o->block == 64 || o->block == 128 || o->block == 256 || o->block == 512 || o->block == 1024;
    if (fixed_block && (!direct || o->sweep_variant != 0)) {
        sh.block = o->block;
        sh.variant = direct ? o->sweep_variant : krylov_group(o);
        return BTE_OK;
    }
    if (!tb && !allow_tune) {
        sh.block = fixed_block ? o->block : 256;
        sh.variant = direct ? (o->sweep_variant ? o->sweep_variant : 2) : krylov_group(o);
        return BTE_OK;
    }
    if (!tb) {
        Shape best;
        int rc = autotune(ctx, o, best);
        if (rc) return rc;
        tb = best.block;
        tv = best.variant;
    }
    sh.block = fixed_block ? o->block : tb;
    sh.variant = (direct && o->sweep_variant != 0) ? o->sweep_variant : tv;
    return BTE_OK;
}

}  // namespace

extern "C" {

// Solves (H + lambda I) y = b for a small symmetric positive semi-definite H by Gaussian
// elimination with partial pivoting; lambda = 1e-12 * max diag keeps it regular.
static void small_solve_ge(std::vector<double> H, std::vector<double> b, int n, std::vector<double>& y) {
    double dmax = 0;
    for (int i = 0; i < n; i++) dmax = std::max(dmax, H[(size_t)i * n + i]);
    for (int i = 0; i < n; i++) H[(size_t)i * n + i] += 1e-12 * dmax + 1e-300;
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++)
            if (std::fabs(H[(size_t)i * n + k]) > std::fabs(H[(size_t)p * n + k])) p = i;
        if (p != k) {
            for (int j = 0; j < n; j++) std::swap(H[(size_t)k * n + j], H[(size_t)p * n + j]);
            std::swap(b[k], b[p]);
        }
        const double piv = H[(size_t)k * n + k];
        for (int i = k + 1; i < n; i++) {
            const double l = H[(size_t)i * n + k] / piv;
            for (int j = k; j < n; j++) H[(size_t)i * n + j] -= l * H[(size_t)k * n + j];
            b[i] -= l * b[k];
        }
    }
    y.assign(n, 0.0);
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int j = i + 1; j < n; j++) s -= H[(size_t)i * n + j] * y[j];
        y[i] = s / H[(size_t)i * n + i];
    }
}

// Factors H + lam I = R^T R (n x n, row-major, upper triangle read) into c, right-looking on the
// upper triangle so that every inner loop runs along a row (independent updates the compiler
// vectorises); false on a pivot that is not positive (H + lam I numerically indefinite).
static bool chol_factor(AaChol& c, const std::vector<double>& H, int n, double lam) {
    constexpr int M = AaChol::M;
    double* U = c.R;
    c.n = -1;
    if (n > M) return false;
    for (int i = 0; i < n; i++) {
        for (int k = i; k < n; k++) U[i * M + k] = H[(size_t)i * n + k];
        U[i * M + i] += lam;
    }
    for (int j = 0; j < n; j++) {
        double* Uj = U + j * M;
        const double d = Uj[j];
        if (!(d > 0) || !std::isfinite(d)) return false;
        const double r = std::sqrt(d), inv = 1.0 / r;
        Uj[j] = r;
        for (int i = j + 1; i < n; i++) Uj[i] *= inv;
        for (int k = j + 1; k < n; k++) {
            const double u = Uj[k];
            double* Uk = U + k * M;
            for (int i = k; i < n; i++) Uk[i] -= u * Uj[i];
        }
    }
    c.n = n;
    c.lam = lam;
    return true;
}

// Drops the first (oldest) column: R without its first column is upper Hessenberg, and Givens
// rotations of neighbouring rows restore the triangle of the remaining columns' factor.
static void chol_drop_first(AaChol& c) {
    constexpr int M = AaChol::M;
    const int n = c.n;
    double* R = c.R;
    for (int i = 0; i < n; i++)
        for (int j = std::max(0, i - 1); j < n - 1; j++) R[i * M + j] = R[i * M + j + 1];
    for (int k = 0; k + 1 < n; k++) {
        double* a = R + k * M;
        double* b = R + (k + 1) * M;
        const double x = a[k], y = b[k], r = std::sqrt(x * x + y * y);  // entries are norms: no overflow
        if (r > 0) {
            const double cs = x / r, sn = y / r;
            for (int j = k; j < n - 1; j++) {
                const double t1 = cs * a[j] + sn * b[j], t2 = cs * b[j] - sn * a[j];
                a[j] = t1;
                b[j] = t2;
            }
        }
        b[k] = 0;
    }
    c.n = n - 1;
}

// Appends a column with inner products h[0, n) with the factored ones and hnn with itself: one
// triangular solve.  False when the new pivot is not safely positive (a column nearly in the span
// of the others), for a new factorisation to decide.
static bool chol_append(AaChol& c, const double* h, double hnn) {
    constexpr int M = AaChol::M;
    const int n = c.n;
    if (n < 0 || n >= M) return false;
    double* R = c.R;
    double ss = 0;
    for (int i = 0; i < n; i++) {
        double s = h[i];
        for (int k = 0; k < i; k++) s -= R[k * M + i] * R[k * M + n];
        const double r = s / R[i * M + i];
        R[i * M + n] = r;
        ss += r * r;
    }
    const double d = hnn + c.lam - ss;
    if (!(d > 1e-8 * (hnn + c.lam)) || !std::isfinite(d)) return false;
    R[n * M + n] = std::sqrt(d);
    c.n = n + 1;
    return true;
}

// R^T z = b, R y = z.
static void chol_solve(const AaChol& c, const std::vector<double>& b, std::vector<double>& y) {
    constexpr int M = AaChol::M;
    const int n = c.n;
    const double* R = c.R;
    std::vector<double> z(b);
    for (int k = 0; k < n; k++) {
        z[k] /= R[k * M + k];
        for (int i = k + 1; i < n; i++) z[i] -= R[k * M + i] * z[k];
    }
    y.assign(n, 0.0);
    for (int i = n - 1; i >= 0; i--) {
        double s = z[i];
        for (int k = i + 1; k < n; k++) s -= R[i * M + k] * y[k];
        y[i] = s / R[i * M + i];
    }
}

// Slot of h_sync that k_signal stores its sequence numbers into.
constexpr int kSyncFlag = 255;

// Waits until the stream has run everything enqueued so far: a one-thread kernel behind that work
// stores a new sequence number into coherent host memory and the host polls it, which ends the wait
// sooner than a stream synchronisation.  The stream is queried every few thousand polls: an error
// ends the wait with that error, and an idle stream whose store is not yet visible falls back to a
// stream synchronisation, so the wait never outlasts the stream's own work.
static int wait_stream(bte_ctx* ctx) {
    const long long seq = ++ctx->sync_seq;
    hipLaunchKernelGGL(k_signal, dim3(1), dim3(64), 0, ctx->stream, seq, (long long*)(ctx->d_sync + kSyncFlag));
    HIPCHK(hipGetLastError());
    const long long* flag = (const long long*)(ctx->h_sync + kSyncFlag);
    for (unsigned spin = 1;; spin++) {
        if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == seq) return BTE_OK;
        if ((spin & 4095u) == 0) {
            const hipError_t e = hipStreamQuery(ctx->stream);
            if (e == hipSuccess) {
                if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == seq) return BTE_OK;
                HIPCHK(hipStreamSynchronize(ctx->stream));
                return BTE_OK;
            }
            if (e != hipErrorNotReady) HIPCHK(e);
        }
        __builtin_ia32_pause();
    }
}

static int aa_reset(bte_ctx* ctx) {
    const int N = ctx->N, F = ctx->F;
    const int m = ctx->aa_plain ? 1 : std::max(1, std::min(ctx->run_opts.accel_memory, 64));
    ctx->aa_m = m;
    if (!ctx->aa_cells_ok) {
        // cells whose previous intensities enter the next right-hand sides: reflecting boundary faces
        std::vector<int> bcells;
        for (int c = 0; c < N; c++) {
            bool refl = false;
            for (int f = 0; f < F; f++) {
                const int nb = ctx->h_nbr[(size_t)c * F + f];
                if (nb >= 0) continue;
                for (size_t k = 0; k < ctx->h_bci.size(); k++)
                    if (ctx->h_bci[k] == nb && ctx->h_bct[k] != 1) refl = true;
            }
            if (refl) bcells.push_back(c);
        }
        ctx->aa_nb = (int)bcells.size();
        std::vector<double> sc(ctx->B);
        for (int b = 0; b < ctx->B; b++) sc[b] = ctx->solid_angle / ctx->h_C[b];
        HIPCHK(upload(ctx->aa_bcells, bcells, ctx->stream));
        HIPCHK(upload(ctx->aa_sc, sc, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));  // sc and bcells are host temporaries
        ctx->aa_cells_ok = true;
    }
    const int S = ctx->B * ctx->D;
    const long long Z = N + (long long)S * ctx->aa_nb;
    ctx->aa_Z = Z;
    for (DBuf* b : {&ctx->aa_x, &ctx->aa_xprev, &ctx->aa_fprev, &ctx->aa_f, &ctx->aa_g, &ctx->aa_dX})
        HIPCHK(b->ensure(sizeof(double) * Z));
    // history columns: only the filled slots [0, aa_order.size()) are ever read
    HIPCHK(ctx->aa_dG.ensure(sizeof(float) * (size_t)m * Z));
    HIPCHK(ctx->aa_dF.ensure(sizeof(float) * (size_t)m * Z));
    if (!ctx->use_scratch) {
        HIPCHK(ctx->aa_part.ensure(sizeof(double) * (size_t)(2 * m + 1) * 64));
        HIPCHK(ctx->aa_dots.ensure(sizeof(double) * (2 * m + 1)));
    }
    HIPCHK(hipMemsetAsync(ctx->aa_x.p, 0, sizeof(double) * Z, ctx->stream));
    ctx->aa_first = true;
    ctx->aa_skip = false;
    ctx->aa_spec = false;
    ctx->aa_steps = 0;
    ctx->aa_last_ff = 0;
    ctx->learn_pending = false;
    ctx->learn_streak = 0;
    ctx->aa_order.clear();
    ctx->aa_H.assign((size_t)m * m, 0.0);
    ctx->aa_chol.n = -1;
    ctx->aa_best = 0;
    return BTE_OK;
}

// The next iteration reads the state aa_x: emission density of its T part, and its boundary
// intensities written into the iterate the next iteration takes as ee_prev.
static int load_state(bte_ctx* ctx) {
    const int N = ctx->N, S = ctx->B * ctx->D, nb = ctx->aa_nb;
    hipLaunchKernelGGL(k_set_density, dim3(grid_for(N, 256)), dim3(256), 0, ctx->stream, N, bands_of(ctx),
                       ctx->aa_x.as<double>(), ctx->dens.as<double>(), ctx->tsrc.as<double>());
    if (nb > 0)
        hipLaunchKernelGGL(k_state_scatter, dim3(grid_for((long long)S * nb, 256)), dim3(256), 0, ctx->stream, N,
                           ctx->D, S, nb, ctx->eei, ctx->aa_bcells.as<int>(), ctx->aa_sc.as<double>(), ctx->aa_x.as<double>(),
                           ctx->ee_final);
    HIPCHK(hipGetLastError());
    return BTE_OK;
}

constexpr int kAaDotG = 2;  // history columns per work-group of the dot products

// The launches of one problem's mixing step, recorded instead of issued: the uneven batch issues
// those of all its problems with one launch per kernel.
struct AaRec {
    bool gather = false, dots = false, update = false;
    AaGather ga{};
    AaDots da{};
    AaUpdate ua{};
    AaCoef coef{};
};

// One Anderson step after an outer iteration produced g = G(x) in ctx->T: forms the next iterate
// x in aa_x and the emission density of x for the following iteration.
// First half of an Anderson step: the state gather, the difference columns and, for a mixing
// step, the dot products, stored into coherent host memory; nothing waits.  Returns in *finish whether
// aa_finish must follow (after the stream has been synchronised).  rec: the launches are recorded
// there instead of issued.
static int aa_launch(bte_ctx* ctx, bool* finish, AaRec* rec = nullptr) {
    const int N = ctx->N, m = ctx->aa_m, S = ctx->B * ctx->D, nb = ctx->aa_nb;
    const int Z = (int)ctx->aa_Z;
    *finish = false;
    ctx->aa_steps++;
    // g = G(x): the new temperature field and reflecting-boundary intensities, gathered in the same
    // pass as the differences f = g - x (and, for a history step, the columns dF and dG and the
    // latest dX)
    // take: a plain step x = g in the same pass (load: with the state's loads for the next iteration)
    auto gather_diff = [&](int first, float* dF, float* dG, int take = 0, bool load = false) {
        AaGather a{N, ctx->D, S, nb, ctx->eei, ctx->aa_bcells.as<int>(), ctx->aa_sc.as<double>(),
                   ctx->T.as<double>(), ctx->ee_final, ctx->aa_g.as<double>(), first, take, ctx->aa_x.as<double>(),
                   ctx->aa_xprev.as<double>(), ctx->aa_fprev.as<double>(), ctx->aa_f.as<double>(),
                   ctx->aa_dX.as<double>(), dF, dG, bands_of(ctx), load ? ctx->dens.as<double>() : (double*)nullptr,
                   ctx->tsrc.as<double>()};
        if (rec) {
            rec->gather = true;
            rec->ga = a;
        } else {
            hipLaunchKernelGGL(k_state_gather_diff, dim3(grid_for(Z, 256)), dim3(256), 0, ctx->stream, a);
        }
    };
    if (ctx->aa_plain) {
        // plain iteration: x_{k+1} = G(x_k); the iteration already left T, density and the boundary
        // intensities of x_{k+1} in place
        gather_diff(1, ctx->aa_dF.as<float>(), ctx->aa_dG.as<float>(), 1);
        HIPCHK(hipGetLastError());
        return BTE_OK;
    }
    // after an accepted correction the step x -> x~ is no iteration step: it enters no history
    // column (the columns stay valid differences of the affine map), the mixing restarts from x~
    const bool skip = ctx->aa_skip;
    ctx->aa_skip = false;
    int slot = 0;
    ctx->aa_popped = false;
    if (!ctx->aa_first && !skip) {
        if ((int)ctx->aa_order.size() < m) {
            slot = (int)ctx->aa_order.size();
        } else {
            slot = ctx->aa_order.front();
            ctx->aa_order.erase(ctx->aa_order.begin());
            ctx->aa_popped = true;
        }
    }
    // the first step is plain, x_1 = G(x_0), taken (with the state's loads) in the same pass
    gather_diff((ctx->aa_first || skip) ? 1 : 0, ctx->aa_dF.as<float>() + (size_t)slot * Z,
                ctx->aa_dG.as<float>() + (size_t)slot * Z, ctx->aa_first ? 1 : 0, ctx->aa_first);
    if (ctx->aa_first) {
        HIPCHK(hipGetLastError());
        ctx->aa_first = false;
        return BTE_OK;
    }
    if (!skip) ctx->aa_order.push_back(slot);
    // the filled slots are [0, n): slots are taken in order and reused only once all m are filled
    const int n = (int)ctx->aa_order.size();
    const int nq = 2 * n + 1;
    const AaDots d{Z, n, ctx->aa_dF.as<float>(), ctx->aa_dF.as<float>() + (size_t)(skip ? 0 : slot) * Z,
                   ctx->aa_f.as<double>(), aa_part_of(ctx), aa_dots_of(ctx), ctx->d_sync + 64};
    if (rec) {
        rec->dots = true;
        rec->da = d;
    } else {
        hipLaunchKernelGGL((k_aa_dots<256, kAaDotG>), dim3(kAaDotBlocks, (n + kAaDotG - 1) / kAaDotG + 1), dim3(256), 0,
                           ctx->stream, d);
        hipLaunchKernelGGL(k_reduce_rows, dim3(std::min(nq, 128)), dim3(256), 0, ctx->stream, nq, kAaDotBlocks,
                           d.partial, d.out, d.host);
    }
    ctx->aa_slot = slot;
    ctx->aa_skip_now = skip;
    ctx->aa_ncols = n;
    *finish = true;
    return BTE_OK;
}

// The least squares of a mixing step, (H + lam I) y = b over the ordered history with
// lam = 1e-12 max diag.  The kept factor follows the step's change of columns (the oldest dropped,
// the newest appended: O(n^2)); it is factorised anew (O(n^3)) when it is not valid, when a new
// pivot is not safely positive, or when lam has moved by more than a factor 100 from the one it
// holds.  A factorisation that fails hands the system to elimination (small_solve_ge).
static void aa_least_squares(bte_ctx* ctx, const std::vector<double>& b, std::vector<double>& y) {
    const int m = ctx->aa_m, n = (int)ctx->aa_order.size();
    AaChol& C = ctx->aa_chol;
    auto Hij = [&](int i, int j) { return ctx->aa_H[(size_t)ctx->aa_order[i] * m + ctx->aa_order[j]]; };
    double dmax = 0;
    for (int i = 0; i < n; i++) dmax = std::max(dmax, Hij(i, i));
    const double lam = 1e-12 * dmax + 1e-300;
    bool ok = C.n >= 0;
    if (ok && !ctx->aa_skip_now) {
        if (ctx->aa_popped && C.n == n) chol_drop_first(C);
        if (C.n == n - 1) {
            double h[AaChol::M];
            for (int i = 0; i + 1 < n; i++) h[i] = Hij(i, n - 1);
            ok = chol_append(C, h, Hij(n - 1, n - 1));
        } else {
            ok = false;
        }
    }
    ok = ok && C.n == n && lam <= 100 * C.lam && 100 * lam >= C.lam;
    if (!ok) {
        std::vector<double> H((size_t)n * n);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) H[(size_t)i * n + j] = Hij(i, j);
        if (!chol_factor(C, H, n, lam)) {
            small_solve_ge(H, b, n, y);
            return;
        }
    }
    chol_solve(C, b, y);
}

// Second half, once the dot products have arrived: the m x m least squares on the host, then the
// update (the coefficients travel as its kernel arguments, or, recorded, with its descriptor) and the
// state.
static int aa_finish(bte_ctx* ctx, AaRec* rec = nullptr) {
    const int m = ctx->aa_m;
    const int Z = (int)ctx->aa_Z;
    const int slot = ctx->aa_slot;
    const int nc = ctx->aa_ncols;  // the dots cover the filled columns [0, nc)
    const double* dots = ctx->h_sync + 64;
    if (!ctx->aa_skip_now)
        for (int j = 0; j < nc; j++) ctx->aa_H[(size_t)slot * m + j] = ctx->aa_H[(size_t)j * m + slot] = dots[j];
    const double ff = dots[2 * nc];
    ctx->aa_last_ff = ff;
    // divergence guard: restart the history when the residual grew far beyond the best seen
    if (ctx->aa_best == 0 || ff < ctx->aa_best) ctx->aa_best = ff;
    AaCoef coef{};
    double* gamma = coef.g;
    if (ff > 1e6 * ctx->aa_best) {
        ctx->aa_order.clear();
        ctx->aa_chol.n = -1;
    } else {
        const int n = (int)ctx->aa_order.size();
        std::vector<double> b(n), y;
        for (int i = 0; i < n; i++) b[i] = dots[nc + ctx->aa_order[i]];
        aa_least_squares(ctx, b, y);
        for (int i = 0; i < n; i++) gamma[ctx->aa_order[i]] = std::isfinite(y[i]) ? y[i] : 0.0;
    }
    // the update and the state's density and boundary intensities in one pass (load_state's work)
    const AaUpdate u{ctx->N, ctx->D, ctx->B * ctx->D, ctx->aa_nb, ctx->eei, nc, ctx->aa_g.as<double>(),
                     ctx->aa_dG.as<float>(), ctx->aa_x.as<double>(), bands_of(ctx), ctx->dens.as<double>(),
                     ctx->aa_bcells.as<int>(), ctx->aa_sc.as<double>(), ctx->ee_final, nullptr,
                     ctx->tsrc.as<double>()};
    if (rec) {
        rec->update = true;
        rec->ua = u;
        rec->coef = coef;
        return BTE_OK;
    }
    hipLaunchKernelGGL(k_aa_update_state, dim3(grid_for(Z, 256)), dim3(256), 0, ctx->stream, u, coef);
    HIPCHK(hipGetLastError());
    return BTE_OK;
}

static int aa_step(bte_ctx* ctx) {
    bool finish = false;
    int rc = aa_launch(ctx, &finish);
    if (rc || !finish) return rc;
    if ((rc = wait_stream(ctx))) return rc;  // the dot products, in coherent host memory
    return aa_finish(ctx);
}

// ---- learned correction ------------------------------------------------------------------------

// Slots of h_sync that the learned correction's norms are stored into: the proposal's residual
// norms {over T, over the whole state} at kSyncLearn, the judge's {residual of the corrected state,
// squared size of the jump} at kSyncJudge; their device copies are red_out_of(ctx)[0, 2) and [2, 4).
constexpr int kSyncLearn = 8, kSyncJudge = 10;

// Two squared norms, sum over [0, n1) of (a1 - b1)^2 and over [0, n2) of (a2 - b2)^2 (a null b counts
// as 0), reduced in a fixed order into red_out_of(ctx)[at, at + 2) and h_sync[kSyncLearn + at, + 2);
// nothing waits.
static int launch_norms2(bte_ctx* ctx, int at, long long n1, const double* a1, const double* b1, long long n2,
                         const double* a2, const double* b2) {
    const int nblk = 32;  // 2 x 32 partial sums: the reduction slice of the scratch
    if (!ctx->use_scratch) {
        HIPCHK(ctx->red_part.ensure(sizeof(double) * 2 * nblk));
        HIPCHK(ctx->red_out.ensure(sizeof(double) * 4));
    }
    hipLaunchKernelGGL(k_sqdiff2<1024>, dim3(nblk), dim3(1024), 0, ctx->stream, n1, a1, b1, n2, a2, b2,
                       red_part_of(ctx));
    hipLaunchKernelGGL(k_reduce_rows, dim3(2), dim3(256), 0, ctx->stream, 2, nblk, red_part_of(ctx),
                       red_out_of(ctx) + at, ctx->d_sync + kSyncLearn + at);
    HIPCHK(hipGetLastError());
    return BTE_OK;
}

// Static node features (host-built once per problem): per cell
//   0 isothermal-wall flag   1 reflecting-wall flag   2 log10(mfp / h) / 2
//   3 (T_interp - T_mid) / T_span, T_interp the inverse-hop-distance blend of the isothermal wall
//     temperatures (linear between two opposite walls)
//   4 d_iso / (d_iso + d_refl) in hops (0 without reflecting walls)
//   5 log10(1 + mean hop distance to the isothermal walls) / 3
//   6 log10(mfp / (h (1 + mean hop distance))) / 2 (a Knudsen number of the whole domain)
//   7 log10(1 + d_iso h / mfp) / 2 (distance from the nearest isothermal wall in mean free paths)
//   8 conductivity fraction of the bands whose mean free path exceeds (d_iso + 1/2) h (the bands
//     for which the cell lies in the Knudsen layer of the nearest isothermal wall)
//   9 conductivity-weighted standard deviation of log10(v_b tau_b) (uniform; band-set spread)
// with mfp = sum_b C v^2 tau / sum_b C v, h the cell size V^(1/dim), d_iso in hops, and band
// conductivity weights k_b = C_b v_b^2 tau_b.
static int learn_prepare(bte_ctx* ctx) {
    if (ctx->learn_stat_ok) return BTE_OK;
    const int N = ctx->N, F = ctx->F, nbc = (int)ctx->h_bci.size();
    double tmin = 1e300, tmax = -1e300;
    std::vector<int> iso;
    for (int k = 0; k < nbc; k++)
        if (ctx->h_bct[k] == 1) {
            iso.push_back(k);
            tmin = std::min(tmin, ctx->h_bcT[k]);
            tmax = std::max(tmax, ctx->h_bcT[k]);
        }
    if (iso.empty()) SETERR(BTE_ERR_BC, "learned correction: the problem has no isothermal wall");
    ctx->learn_Tmid = 0.5 * (tmin + tmax);
    ctx->learn_span = tmax > tmin ? tmax - tmin : 1.0;
    const int INF = 1 << 29;
    // multi-source breadth-first search over interior faces from the cells of the selected walls
    auto bfs = [&](auto&& seed_bc, std::vector<int>& d) {
        d.assign(N, INF);
        std::vector<int> q;
        q.reserve(N);
        for (int c = 0; c < N; c++)
            for (int f = 0; f < F; f++) {
                const int nb = ctx->h_nbr[(size_t)c * F + f];
                if (nb < 0 && seed_bc(nb)) {
                    d[c] = 0;
                    q.push_back(c);
                    break;
                }
            }
        for (size_t h = 0; h < q.size(); h++) {
            const int c = q[h];
            for (int f = 0; f < F; f++) {
                const int nb = ctx->h_nbr[(size_t)c * F + f];
                if (nb >= 0 && d[nb] == INF) {
                    d[nb] = d[c] + 1;
                    q.push_back(nb);
                }
            }
        }
    };
    std::vector<std::vector<int>> d_iso(iso.size());
    for (size_t i = 0; i < iso.size(); i++) {
        const int idx = ctx->h_bci[iso[i]];
        bfs([idx](int nb) { return nb == idx; }, d_iso[i]);
    }
    std::vector<int> d_refl;
    bool any_refl = false;
    bfs(
        [&](int nb) {
            for (int k = 0; k < nbc; k++)
                if (ctx->h_bci[k] == nb && ctx->h_bct[k] != 1) return any_refl = true;
            return false;
        },
        d_refl);
    double num = 0, den = 0;
    for (int b = 0; b < ctx->B; b++) {
        num += ctx->h_C[b] * ctx->h_v[b] * ctx->h_v[b] * ctx->h_tau[b];
        den += ctx->h_C[b] * ctx->h_v[b];
    }
    const double mfp = den > 0 ? num / den : 0;
    double lm = 0, lm2 = 0;
    for (int b = 0; b < ctx->B; b++) {
        const double kb = ctx->h_C[b] * ctx->h_v[b] * ctx->h_v[b] * ctx->h_tau[b] / num;
        const double l = std::log10(ctx->h_v[b] * ctx->h_tau[b]);
        lm += kb * l;
        lm2 += kb * l * l;
    }
    const float spread = (float)std::sqrt(std::max(0.0, lm2 - lm * lm));
    std::vector<float> st((size_t)N * kLearnStatic);
    for (int c = 0; c < N; c++) {
        double wsum = 0, tsum = 0, dsum = 0;
        int nreach = 0, dmin = INF;
        for (size_t i = 0; i < iso.size(); i++) {
            const int dk = d_iso[i][c];
            if (dk >= INF) continue;
            const double w = 1.0 / (dk + 0.5);
            wsum += w;
            tsum += w * ctx->h_bcT[iso[i]];
            dsum += dk;
            nreach++;
            dmin = std::min(dmin, dk);
        }
        const double t_int = wsum > 0 ? tsum / wsum : ctx->learn_Tmid;
        const double dmean = nreach ? dsum / nreach : 0.0;
        const double vol = std::max(ctx->h_vol[c], 1e-300);
        const double h = ctx->dim == 1 ? vol : (ctx->dim == 2 ? std::sqrt(vol) : std::cbrt(vol));
        const double kn = std::max(mfp / h, 1e-12);
        float* o = st.data() + (size_t)c * kLearnStatic;
        o[0] = dmin == 0 ? 1.f : 0.f;
        o[1] = d_refl[c] == 0 ? 1.f : 0.f;
        o[2] = (float)(std::log10(kn) / 2);
        o[3] = (float)((t_int - ctx->learn_Tmid) / ctx->learn_span);
        o[4] = (any_refl && d_refl[c] < INF && dmin < INF) ? (float)((double)dmin / (dmin + d_refl[c] + 1e-30)) : 0.f;
        o[5] = (float)(std::log10(1.0 + dmean) / 3);
        o[6] = (float)(std::log10(kn / (1.0 + dmean)) / 2);
        o[7] = dmin < INF ? (float)(std::log10(1.0 + dmin / kn) / 2) : 4.f;
        double bal = 0;
        for (int b = 0; b < ctx->B; b++)
            if (ctx->h_v[b] * ctx->h_tau[b] > (dmin + 0.5) * h)
                bal += ctx->h_C[b] * ctx->h_v[b] * ctx->h_v[b] * ctx->h_tau[b] / num;
        o[8] = (float)bal;
        o[9] = spread;
    }
    HIPCHK(upload(ctx->learn_stat, st, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));  // st is a host temporary
    ctx->learn_stat_ok = true;
    return BTE_OK;
}

// Features of the current state into X (device, [N][16]), with the residual norms of the last
// evaluated state (over the temperature part and over the whole state) formed first on the device:
// the features scale by the first, and learn_judge reads both.  Nothing waits.
static int learn_features(bte_ctx* ctx, float* X) {
    int rc;
    if ((rc = learn_prepare(ctx))) return rc;
    const int N = ctx->N;
    const long long Z = ctx->aa_Z;
    if ((rc = launch_norms2(ctx, 0, N, ctx->aa_f.as<double>(), nullptr, Z, ctx->aa_f.as<double>(), nullptr)))
        return rc;
    const double* dXl = nullptr;
    const float* dFl = nullptr;
    if (!ctx->aa_plain && !ctx->aa_order.empty()) {
        const int slot = ctx->aa_order.back();
        dXl = ctx->aa_dX.as<double>();  // the step of that slot
        dFl = ctx->aa_dF.as<float>() + (size_t)slot * Z;
    }
    hipLaunchKernelGGL(k_learn_features, dim3(grid_for(N, 256)), dim3(256), 0, ctx->stream, N, ctx->aa_f.as<double>(),
                       ctx->aa_x.as<double>(), ctx->aa_xprev.as<double>(), dXl, dFl, ctx->learn_stat.as<float>(),
                       red_out_of(ctx), ctx->learn_Tmid, ctx->learn_span, X);
    HIPCHK(hipGetLastError());
    return BTE_OK;
}

// Points the attached network's graph at this context's mesh (a network may serve several contexts).
static int learn_graph(bte_ctx* ctx) {
    bte_gnn* g = ctx->gnn;
    if (g->graph_tag == ctx->geo_tag && g->N == ctx->N) return BTE_OK;
    const int rc = bte_gnn_set_graph(g, ctx->N, ctx->F, ctx->h_nbr.data());
    if (rc) SETERR(rc, std::string("learned correction: ") + g->err);
    g->graph_tag = ctx->geo_tag;
    return BTE_OK;
}

// Proposes a correction of the next state: x_next += T_span * GNN(features).  The next outer
// iteration runs from the corrected state and learn_judge decides whether it stands.  Nothing
// waits: a state with no residual to correct (learn_judge finds it from the norms) is left as it is
// by k_apply_correction, and the proposal then counts as not made.
static int learn_propose(bte_ctx* ctx) {
    GnnNet* net = ctx->gnn;
    const int N = ctx->N;
    int rc;
    if ((rc = learn_graph(ctx))) return rc;
    if ((rc = learn_features(ctx, net->X.as<float>()))) return rc;
    if (gnn_infer_device(net, ctx->stream)) SETERR(BTE_ERR_HIP, "learned correction: network launch failed");
    const long long Z = ctx->aa_Z;
    HIPCHK(ctx->learn_backup.ensure(sizeof(double) * Z));
    // backup, correction and the corrected state's loads (load_state's work) in one pass
    hipLaunchKernelGGL(k_apply_correction, dim3(grid_for(Z, 256)), dim3(256), 0, ctx->stream, N, ctx->D,
                       ctx->B * ctx->D, ctx->aa_nb, ctx->eei, ctx->aa_bcells.as<int>(), net->y.as<float>(), ctx->learn_span,
                       red_out_of(ctx), ctx->aa_x.as<double>(), ctx->learn_backup.as<double>(), bands_of(ctx),
                       ctx->dens.as<double>(), ctx->tsrc.as<double>(), ctx->aa_sc.as<double>(), ctx->ee_final);
    HIPCHK(hipGetLastError());
    ctx->learn_pending = true;
    ctx->run_proposed++;
    return BTE_OK;
}

// The judge's inputs for the iteration that ran from a corrected state: the squared norms of its
// residual G(x~) - x~ over the state (G gathered in the same pass) and of the jump x~ - x over the
// temperature part, into red_out_of(ctx)[2, 4) and h_sync[kSyncJudge, + 2); nothing waits.
static int launch_judge(bte_ctx* ctx) {
    const int N = ctx->N, S = ctx->B * ctx->D, nb = ctx->aa_nb, nblk = 32;
    if (!ctx->use_scratch) {
        HIPCHK(ctx->red_part.ensure(sizeof(double) * 2 * nblk));
        HIPCHK(ctx->red_out.ensure(sizeof(double) * 4));
    }
    hipLaunchKernelGGL(k_judge_norms<1024>, dim3(nblk), dim3(1024), 0, ctx->stream, N, ctx->D, S, nb, ctx->eei,
                       ctx->aa_bcells.as<int>(), ctx->aa_sc.as<double>(), ctx->T.as<double>(), ctx->ee_final,
                       ctx->aa_x.as<double>(), ctx->learn_backup.as<double>(), red_part_of(ctx));
    hipLaunchKernelGGL(k_reduce_rows, dim3(2), dim3(256), 0, ctx->stream, 2, nblk, red_part_of(ctx),
                       red_out_of(ctx) + 2, ctx->d_sync + kSyncJudge);
    HIPCHK(hipGetLastError());
    return BTE_OK;
}

// After the outer iteration from a corrected state: accept when its residual |G(x~) - x~| does
// not exceed the residual of the state evaluated before the correction.
// void_: the state had no residual to correct, so no correction was applied and the iteration ran
// from the uncorrected state.  The judge's norms come from launch_judge behind the iteration
// (judge_spec), or are formed here.
static int learn_judge(bte_ctx* ctx, bool& accepted, bool& void_) {
    int rc;
    accepted = void_ = false;
    if (!ctx->judge_spec && ((rc = launch_judge(ctx)) || (rc = wait_stream(ctx)))) return rc;
    ctx->judge_spec = false;
    const double ffT = ctx->h_sync[kSyncLearn], ffZ = ctx->h_sync[kSyncLearn + 1];
    if (!(ffT > 0) || !std::isfinite(ffZ)) {
        void_ = true;
        return BTE_OK;
    }
    ctx->learn_ref_ff = ffZ;
    const double ff = ctx->h_sync[kSyncJudge];
    ctx->learn_jump = ctx->h_sync[kSyncJudge + 1];
    accepted = std::isfinite(ff) && ff <= ctx->learn_ref_ff;
    ctx->learn_ratio = ctx->learn_ref_ff > 0 ? std::sqrt(ff / ctx->learn_ref_ff) : 0.0;
    return BTE_OK;
}

// Rejected correction: restore the uncorrected state and undo the outer iteration that ran from it
// (its input intensities are untouched; the iteration only wrote the other buffer).
static int learn_undo(bte_ctx* ctx) {
    HIPCHK(hipMemcpyAsync(ctx->aa_x.p, ctx->learn_backup.p, sizeof(double) * ctx->aa_Z, hipMemcpyDeviceToDevice,
                          ctx->stream));
    ctx->run_iter--;
    ctx->ee_final = (ctx->run_iter & 1) ? ctx->eeB.as<double>() : ctx->eeA.as<double>();
    return load_state(ctx);
}

static int check_opts(bte_ctx* ctx, const bte_opts& o) {
    if (o.precision != BTE_PREC_FP64 && o.precision != BTE_PREC_MIXED) SETERR(BTE_ERR_ARG, "unknown precision");
    if (o.precond != BTE_PC_SWEEP && o.precond != BTE_PC_JACOBI) SETERR(BTE_ERR_ARG, "unknown preconditioner");
    if (o.inner_maxit <= 0 || o.inner_tol <= 0) SETERR(BTE_ERR_ARG, "inner_maxit and inner_tol must be positive");
    if (o.accel != BTE_ACCEL_NONE && o.accel != BTE_ACCEL_ANDERSON) SETERR(BTE_ERR_ARG, "unknown outer acceleration");
    if (o.intensity_layout < 0 || o.intensity_layout > 2) SETERR(BTE_ERR_ARG, "unknown intensity layout");
    if (o.learned) {
        if (o.learned_every < 0) SETERR(BTE_ERR_ARG, "learned_every must not be negative");
        if (o.learned_every > 0 && !ctx->gnn)
            SETERR(BTE_ERR_STATE, "learned correction requested but no network is attached");
        if (ctx->gnn && ctx->gnn->device != ctx->device)
            SETERR(BTE_ERR_ARG, "the attached network lives on another device");
    }
    return BTE_OK;
}

// Whether the captured iteration graphs were recorded with the options and launch shape that
// enqueue_iteration would use now: the solver form and every SolveCfg field baked into the launches.
static bool same_capture(const bte_ctx* ctx, const bte_opts& o, const Shape& sh) {
    const bte_opts& g = ctx->graph_opts;
    return ctx->graph_valid && ctx->graph[0] && ctx->graph[1] && ctx->graph_block == sh.block &&
           ctx->graph_variant == sh.variant && g.precision == o.precision && g.precond == o.precond &&
           (g.direct_sweep != 0) == (o.direct_sweep != 0) && g.inner_maxit == o.inner_maxit &&
           g.inner_tol == o.inner_tol && std::max(1, g.refine_steps) == std::max(1, o.refine_steps) &&
           g.inner_tol32 == o.inner_tol32 && ctx->graph_inter == ctx->ee_inter;
}

int bte_begin(bte_ctx* ctx, const bte_opts* opts) {
    if (!ctx) return BTE_ERR_ARG;
    bte_opts o;
    if (opts)
        o = *opts;
    else
        bte_default_opts(&o);
    int rc = check_opts(ctx, o);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const bool timing = o.timing != 0;
    auto t0 = std::chrono::steady_clock::now();
    if ((rc = prepare(ctx, timing))) return rc;
    if (o.precision == BTE_PREC_MIXED && (rc = ensure_fp32(ctx))) return rc;
    if ((rc = ensure_work(ctx, o.precision))) return rc;
    const int N = ctx->N, B = ctx->B, D = ctx->D, S = B * D;
    const size_t SN = (size_t)S * N;
    Shape shape;
    // the autotuner creates events and overwrites state; bte_solve_noalloc takes a tuned or the
    // default shape instead
    if ((rc = pick_shape(ctx, &o, shape, !ctx->use_scratch))) return rc;
    ctx->block_used = shape.block;
    // the sweep variant that actually runs (a requested form the mesh cannot use falls back to 2)
    ctx->variant_used = !direct_path(ctx, &o)                              ? 0
                        : (shape.variant == 6 && ctx->B >= 2 && ring_size(ctx, nullptr, 2) > 0) ? 6
                        : (shape.variant == 7 && ctx->B >= 4 && ring_size(ctx, nullptr, 4) > 0) ? 7
                        : (shape.variant == 5 && ring_size(ctx) > 0)      ? 5
                        : (shape.variant == 4 && ring_size(ctx) > 0)      ? 4
                        : (shape.variant == 8 && ctx->ring_r2 > 0)        ? 8
                        : (shape.variant == 3 && lds_fits(ctx))           ? 3
                        : (shape.variant == 1)                            ? 1
                                                                          : 2;
    ctx->group_used = direct_path(ctx, &o) ? 0
                      : group_ok(ctx, &o, shape.variant, shape.block) ? shape.variant
                                                                       : 1;
    // intensity layout: each cell's bands interleaved for the pipelined sweep when asked, or by
    // default in 3D with 4 or more bands (the co-scheduled bands of a direction then write a
    // cell's intensities together, into whole sectors); band-major for every other form
    set_layout(ctx, direct_path(ctx, &o) && (ctx->variant_used == 4 || ctx->variant_used == 8) && ctx->B >= 2 &&
                        (o.intensity_layout == 2 || (o.intensity_layout == 0 && ctx->dim == 3 && ctx->B >= 4)));
    ctx->run_inter = ctx->ee_inter;
    // the zero initial intensities (the first iteration's ee_prev); the buffer it writes is cleared
    // only for the Krylov forms, whose initial guess it is (a sweep writes every cell)
    HIPCHK(hipMemsetAsync(ctx->eeA.p, 0, ctx->eeA.bytes, ctx->stream));
    if (!direct_path(ctx, &o)) HIPCHK(hipMemsetAsync(ctx->eeB.p, 0, sizeof(double) * SN, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->dens.p, 0, sizeof(double) * (size_t)B * N, ctx->stream));
    hipLaunchKernelGGL(k_fill, dim3(grid_for(N, 256)), dim3(256), 0, ctx->stream, N, ctx->T_ref,
                       ctx->tsrc.as<double>());  // the zero density's source temperature
    ctx->tsrc_ok = true;
    HIPCHK(hipMemsetAsync(ctx->T.p, 0, sizeof(double) * N, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->iters.p, 0, sizeof(int) * S, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->fails.p, 0, sizeof(int) * S, ctx->stream));
    double* buf[2] = {ctx->eeA.as<double>(), ctx->eeB.as<double>()};
    const bool use_graph = o.use_graph && !timing;
    if (use_graph && !same_capture(ctx, o, shape)) {
        HIPCHK(hipStreamSynchronize(ctx->stream));  // capture starts on an idle stream
        drop_graphs(ctx);
        for (int par = 0; par < 2; par++) {
            hipGraph_t g = nullptr;
            HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_iteration(ctx, &o, shape, buf[par], buf[1 - par], false);
            hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
            if (rc) return rc;
            HIPCHK(ec);
            HIPCHK(hipGraphInstantiate(&ctx->graph[par], g, nullptr, nullptr, 0));
            HIPCHK(hipGraphDestroy(g));
        }
        ctx->graph_valid = true;
        ctx->graph_inter = ctx->ee_inter;
        ctx->graph_opts = o;
        ctx->graph_block = shape.block;
        ctx->graph_variant = shape.variant;
    }
    ctx->run_opts = o;
    ctx->aa_plain = o.accel != BTE_ACCEL_ANDERSON;
    if (o.accel == BTE_ACCEL_ANDERSON || o.learned) {
        if ((rc = aa_reset(ctx))) return rc;
        // x_0 = T_ref everywhere: the zero emission density of the first iteration
        hipLaunchKernelGGL(k_fill, dim3(grid_for(N, 256)), dim3(256), 0, ctx->stream, N, ctx->T_ref,
                           ctx->aa_x.as<double>());
        HIPCHK(hipGetLastError());
    }
    if (o.learned) {
        if ((rc = learn_prepare(ctx))) return rc;
        if (ctx->gnn && (rc = learn_graph(ctx))) return rc;
    }
    ctx->run_accepted = ctx->run_proposed = ctx->run_rejected = 0;
    ctx->judge_spec = false;
    ctx->run_shape_block = shape.block;
    ctx->run_shape_variant = shape.variant;
    ctx->run_graph = use_graph;
    ctx->run_iter = 0;
    ctx->run_converged = false;
    ctx->run_active = true;
    ctx->run_margins.clear();
    ctx->run_solve_seconds = 0;
    ctx->ee_final = nullptr;
    ctx->run_setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (o.verbose)
        std::printf("N_cell: %d\nN_dir: %d\nN_band: %d\nnum_theta: %d\nnum_phi: %d\n\n", N, D, B, ctx->num_theta,
                    ctx->num_phi);
    return BTE_OK;
}

// Runs one plain outer iteration (diffuse data, inner solves, T update, margin) and returns the margin.
static int run_iteration(bte_ctx* ctx, double& m) {
    const bte_opts& o = ctx->run_opts;
    double* buf[2] = {ctx->eeA.as<double>(), ctx->eeB.as<double>()};
    const int par = ctx->run_iter & 1;  // iteration parity: ee_prev = buf[par], ee_curr = buf[1-par]
    if (ctx->run_graph) {
        if (!ctx->graph[par]) SETERR(BTE_ERR_STATE, "iteration state was invalidated (inputs changed); call bte_begin");
        HIPCHK(hipGraphLaunch(ctx->graph[par], ctx->stream));
    } else {
        Shape sh;
        sh.block = ctx->run_shape_block;
        sh.variant = ctx->run_shape_variant;
        int rc = enqueue_iteration(ctx, &o, sh, buf[par], buf[1 - par], o.timing != 0);
        if (rc) return rc;
    }
    ctx->ee_final = buf[1 - par];
    // A mixing step reads the new state whatever the margin: its gather and dot products follow the
    // margin's copy, so that one synchronisation brings both.  Not when a learned correction awaits
    // its judgement (which may undo the iteration) nor for the first state step.
    ctx->aa_spec = false;
    if (o.accel == BTE_ACCEL_ANDERSON && !ctx->aa_plain && !ctx->aa_first && !ctx->learn_pending && ctx->aa_Z > 0) {
        ctx->aa_saved_steps = ctx->aa_steps;
        ctx->aa_saved_skip = ctx->aa_skip;
        ctx->aa_saved_order = ctx->aa_order;
        int rc = aa_launch(ctx, &ctx->aa_spec_finish);
        if (rc) return rc;
        ctx->aa_spec = true;
    }
    // likewise the judge's norms when the iteration ran from a corrected state
    ctx->judge_spec = false;
    if (ctx->learn_pending && ctx->aa_Z > 0) {
        if (int rc = launch_judge(ctx)) return rc;
        ctx->judge_spec = true;
    }
    if (int rc = wait_stream(ctx)) return rc;
    m = ctx->h_sync[0];  // stored by the iteration's margin kernel
    ctx->run_iter++;
    return BTE_OK;
}

// Bookkeeping after an outer iteration of ctx produced margin m (ee_final, run_iter already
// advanced): judge a pending learned correction (a rejected one is undone and the iteration does not
// count), record the margin, test convergence, then mix the state and propose the next correction.
// The parts before and after the state mixing are separate (after_judge, after_mix) so that the
// uneven batch can issue the mixing steps of all its problems with one synchronisation.
static int after_judge(bte_ctx* ctx, double m, bool& counted, bool& mix) {
    const bte_opts& o = ctx->run_opts;
    int rc;
    counted = false;
    mix = false;
    if (ctx->learn_pending) {
        ctx->learn_pending = false;
        bool ok = false, void_ = false;
        if ((rc = learn_judge(ctx, ok, void_))) return rc;
        if (o.verbose && !void_)
            std::printf("learned correction %s (residual ratio %.4g)\n", ok ? "accepted" : "rejected", ctx->learn_ratio);
        if (void_) {
            ctx->run_proposed--;  // nothing was corrected: the iteration stands as a plain one
        } else if (ok) {
            ctx->run_accepted++;
            ctx->learn_streak = 0;
            // the jump to the corrected state enters no mixing column: a least-squares step over it
            // would take the correction back (the residual hardly sees smooth errors)
            ctx->aa_skip = ctx->learn_jump > 0;
        } else {
            // run the iteration again from the uncorrected state
            ctx->run_rejected++;
            ctx->learn_streak++;
            return learn_undo(ctx);
        }
    }
    counted = true;
    ctx->run_margins.push_back(m);
    if (o.verbose) {
        std::printf("----------------------------------------------------------------------------------\n");
        std::printf("Iteration #%d\t Margin per band per cell: %g\n", ctx->run_iter - 1, m);
    }
    if (m <= o.stop_margin) {
        ctx->run_converged = true;
        if (o.verbose) std::printf("\n");
    } else if (o.accel == BTE_ACCEL_ANDERSON || o.learned) {
        mix = true;
    }
    return BTE_OK;
}

// After the mixing step: corrections at state steps 1, 1 + every, 1 + 2 every, ... until two in a
// row fail.
static int after_mix(bte_ctx* ctx) {
    const bte_opts& o = ctx->run_opts;
    if (o.learned && o.learned_every > 0 && ctx->learn_streak < 2 && (ctx->aa_steps - 1) % o.learned_every == 0)
        return learn_propose(ctx);
    return BTE_OK;
}

static int after_iteration(bte_ctx* ctx, double m, bool& counted) {
    bool mix = false;
    int rc = after_judge(ctx, m, counted, mix);
    if (rc) return rc;
    if (ctx->aa_spec) {  // the step's first half was launched with the iteration
        ctx->aa_spec = false;
        if (!mix) {  // the margin ended the solve: the step is not taken
            ctx->aa_steps = ctx->aa_saved_steps;
            ctx->aa_skip = ctx->aa_saved_skip;
            ctx->aa_order = ctx->aa_saved_order;
            return BTE_OK;
        }
        if (ctx->aa_spec_finish && (rc = aa_finish(ctx))) return rc;
        return after_mix(ctx);
    }
    if (!mix) return BTE_OK;
    if ((rc = aa_step(ctx))) return rc;
    return after_mix(ctx);
}

int bte_step(bte_ctx* ctx, int n_iter, int* n_done, int* converged) {
    if (!ctx) return BTE_ERR_ARG;
    if (!ctx->run_active) SETERR(BTE_ERR_STATE, "bte_step: call bte_begin first");
    HIPCHK(hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();
    int done = 0;
    while (done < n_iter && !ctx->run_converged) {
        double m = 0;
        int rc = run_iteration(ctx, m);
        if (rc) return rc;
        bool counted = false;
        if ((rc = after_iteration(ctx, m, counted))) return rc;
        if (counted) done++;
    }
    ctx->run_solve_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (n_done) *n_done = done;
    if (converged) *converged = ctx->run_converged ? 1 : 0;
    return BTE_OK;
}

int bte_attach_gnn(bte_ctx* ctx, bte_gnn* g) {
    if (!ctx) return BTE_ERR_ARG;
    if (g && g->device != ctx->device) SETERR(BTE_ERR_ARG, "bte_attach_gnn: the network lives on another device");
    ctx->gnn = g;
    ctx->learn_pending = false;
    return BTE_OK;
}

int bte_learn_features(bte_ctx* ctx, float* X, double* x_next, double* span) {
    if (!ctx || !X) return BTE_ERR_ARG;
    if (!ctx->run_active || ctx->aa_Z == 0 || ctx->aa_steps == 0 ||
        !(ctx->run_opts.accel == BTE_ACCEL_ANDERSON || ctx->run_opts.learned))
        SETERR(BTE_ERR_STATE, "bte_learn_features: needs an iteration begun with accel or learned and one bte_step");
    HIPCHK(hipSetDevice(ctx->device));
    const int N = ctx->N;
    HIPCHK(ctx->learn_X.ensure(sizeof(float) * (size_t)N * kGnnIn));
    int rc;
    if ((rc = learn_features(ctx, ctx->learn_X.as<float>()))) return rc;
    std::vector<float> hX((size_t)N * kGnnIn);
    std::vector<double> hx(N);
    HIPCHK(hipMemcpyAsync(hX.data(), ctx->learn_X.p, sizeof(float) * hX.size(), hipMemcpyDeviceToHost, ctx->stream));
    if (x_next) HIPCHK(hipMemcpyAsync(hx.data(), ctx->aa_x.p, sizeof(double) * N, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    to_external(ctx, hX.data(), X, kGnnIn);
    if (x_next) to_external(ctx, hx.data(), x_next);
    if (span) *span = ctx->learn_span;
    return BTE_OK;
}

int bte_get_temperature(bte_ctx* ctx, double* T, double* T_band) {
    if (!ctx) return BTE_ERR_ARG;
    if (!ctx->run_active) SETERR(BTE_ERR_STATE, "no iteration state");
    HIPCHK(hipSetDevice(ctx->device));
    const int N = ctx->N, B = ctx->B;
    std::vector<double> hT, hTb;
    if (T) {
        hT.resize(N);
        HIPCHK(hipMemcpyAsync(hT.data(), ctx->T.p, sizeof(double) * N, hipMemcpyDeviceToHost, ctx->stream));
    }
    if (T_band) {
        hTb.resize((size_t)B * N);
        HIPCHK(hipMemcpyAsync(hTb.data(), ctx->Tband.p, sizeof(double) * B * N, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (T) to_external(ctx, hT.data(), T);
    if (T_band)
        for (int i = 0; i < N; i++)
            for (int b = 0; b < B; b++) T_band[(size_t)ext_cell(ctx, i) * B + b] = hTb[(size_t)b * N + i];
    return BTE_OK;
}

int bte_set_temperature(bte_ctx* ctx, const double* T) {
    if (!ctx || !T) return BTE_ERR_ARG;
    if (!ctx->run_active) SETERR(BTE_ERR_STATE, "no iteration state");
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<double> hT(ctx->N);
    to_internal(ctx, T, hT.data());
    HIPCHK(hipMemcpyAsync(ctx->T.p, hT.data(), sizeof(double) * ctx->N, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_set_density, dim3(grid_for(ctx->N, 256)), dim3(256), 0, ctx->stream, ctx->N, bands_of(ctx),
                       ctx->T.as<double>(), ctx->dens.as<double>(), ctx->tsrc.as<double>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));  // hT is a host temporary
    return BTE_OK;
}

int bte_finish(bte_ctx* ctx, bte_result* res) {
    if (!ctx || !res) return BTE_ERR_ARG;
    if (!ctx->run_active) SETERR(BTE_ERR_STATE, "bte_finish: call bte_begin first");
    HIPCHK(hipSetDevice(ctx->device));
    const bte_opts& o = ctx->run_opts;
    const int N = ctx->N, B = ctx->B, S = B * ctx->D;
    const int it = ctx->run_iter;
    res->setup_seconds = ctx->run_setup_seconds;
    res->solve_seconds = ctx->run_solve_seconds;
    res->iterations = it;
    res->converged = ctx->run_converged ? 1 : 0;
    res->learned_accepted = ctx->run_accepted;
    res->learned_proposed = ctx->run_proposed;
    if (res->margins)
        for (int i = 0; i < it && i < res->margins_capacity; i++) res->margins[i] = ctx->run_margins[i];
    // heat flux of the last iterate
    const int nbc = (int)ctx->h_bci.size();
    if (nbc > 0 && it > 0) {
        Timer tm(ctx, KA_FLUX, o.timing != 0);
        hipLaunchKernelGGL(k_flux<256>, dim3(nbc * B), dim3(256), 0, ctx->stream, geo_of(ctx), quad_of(ctx), bnd_of(ctx),
                           bands_of(ctx), ctx->DM, ctx->ee_final, bflux_of(ctx));
        HIPCHK(hipGetLastError());
    }
    std::vector<int> hit(S), hfail(S);
    HIPCHK(hipMemcpyAsync(hit.data(), ctx->iters.p, sizeof(int) * S, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(hfail.data(), ctx->fails.p, sizeof(int) * S, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<double> hT(N), hTb((size_t)B * N), hbf((size_t)nbc * B);
    HIPCHK(hipMemcpyAsync(hT.data(), ctx->T.p, sizeof(double) * N, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(hTb.data(), ctx->Tband.p, sizeof(double) * B * N, hipMemcpyDeviceToHost, ctx->stream));
    if (nbc > 0 && it > 0)
        HIPCHK(hipMemcpyAsync(hbf.data(), bflux_of(ctx), sizeof(double) * nbc * B, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // a proposal still awaiting its judgement when the solve stopped: with no residual to correct
    // (its norms have arrived with the synchronisation) it was not made
    if (ctx->learn_pending) {
        ctx->learn_pending = false;
        if (!(ctx->h_sync[kSyncLearn] > 0) || !std::isfinite(ctx->h_sync[kSyncLearn + 1])) ctx->run_proposed--;
        res->learned_proposed = ctx->run_proposed;
    }
    res->inner_iters = 0;
    res->inner_failures = 0;
    for (int s = 0; s < S; s++) {
        res->inner_iters += hit[s];
        res->inner_failures += hfail[s];
    }
    if (res->T) to_external(ctx, hT.data(), res->T);
    if (res->T_band)
        for (int i = 0; i < N; i++)
            for (int b = 0; b < B; b++) res->T_band[(size_t)ext_cell(ctx, i) * B + b] = hTb[(size_t)b * N + i];
    if (res->bc_band_flux) std::copy(hbf.begin(), hbf.end(), res->bc_band_flux);
    if (res->bc_flux)
        for (int k = 0; k < nbc; k++) {
            double s = 0;
            for (int b = 0; b < B; b++) s += hbf[(size_t)k * B + b];
            res->bc_flux[k] = s;
        }
    ctx->spmv_rows = res->inner_iters * 2LL * N;
    if (o.verbose)
        for (int k = 0; k < nbc; k++) {
            double s = 0;
            for (int b = 0; b < B; b++) s += hbf[(size_t)k * B + b];
            std::printf("heat flux of Boundary Condition #%d: %g\n", -1 - ctx->h_bci[k], s);
        }
    ctx->run_active = false;
    if (res->inner_failures > 0) {
        ctx->err = "some inner solves reached inner_maxit";
        return BTE_ERR_NOCONV;
    }
    return BTE_OK;
}

int bte_solve(bte_ctx* ctx, const bte_opts* opts, bte_result* res) {
    if (!ctx || !res) return BTE_ERR_ARG;
    bte_opts o;
    if (opts)
        o = *opts;
    else
        bte_default_opts(&o);
    int rc = bte_begin(ctx, &o);
    if (rc) return rc;
    if ((rc = bte_step(ctx, o.max_iter, nullptr, nullptr))) return rc;
    return bte_finish(ctx, res);
}

// bte_solve without device allocations: the reductions (margin partial sums and margin, squared
// norms, Anderson dot products, heat flux staging) use this device's scratch slices, graphs are
// neither captured nor replayed (instantiation may allocate), the launch shape is a tuned or the
// default one, and every other device buffer must already be large enough -- from an earlier solve
// of this form with inputs at least this large.  If one is not, the solve stops with
// BTE_ERR_STATE before any work depends on it.
int bte_solve_noalloc(bte_ctx* ctx, const bte_opts* opts, bte_result* res) {
    if (!ctx || !res) return BTE_ERR_ARG;
    if (!ctx->scratch_dev)
        SETERR(BTE_ERR_HIP, std::string("bte_solve_noalloc: the device's scratch array could not be resolved: ") +
                                hipGetErrorString(ctx->scratch_rc));
    bte_opts o;
    if (opts)
        o = *opts;
    else
        bte_default_opts(&o);
    o.use_graph = 0;
    o.timing = 0;
    const size_t flux = ctx->h_bci.size() * (size_t)std::max(ctx->B, 1);
    if (flux > kScrFluxMax)
        SETERR(BTE_ERR_ARG, "bte_solve_noalloc: " + std::to_string(flux) + " (boundary condition, band) pairs exceed "
                                "the scratch's " + std::to_string(kScrFluxMax));
    struct Mode {
        bte_ctx* c;
        explicit Mode(bte_ctx* x) : c(x) {
            c->use_scratch = true;
            AllocGuard::on() = true;
            AllocGuard::tripped() = false;
        }
        ~Mode() {
            c->use_scratch = false;
            AllocGuard::on() = false;
        }
    } mode(ctx);
    int rc = bte_begin(ctx, &o);
    if (!rc) rc = bte_step(ctx, o.max_iter, nullptr, nullptr);
    if (!rc) rc = bte_finish(ctx, res);
    if (AllocGuard::tripped()) {
        ctx->run_active = false;
        SETERR(BTE_ERR_STATE, "bte_solve_noalloc: the context's device buffers are smaller than this solve needs; "
                              "run bte_solve once with this form and inputs at least this large");
    }
    return rc;
}

int bte_get_intensity(bte_ctx* ctx, double* ee) {
    if (!ctx || !ee) return BTE_ERR_ARG;
    if (!ctx->ee_final) SETERR(BTE_ERR_STATE, "no solve has run");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t S = (size_t)ctx->B * ctx->D, N = ctx->N;
    std::vector<double> h(S * N);
    if (ctx->run_inter) {  // to band-major
        const EeIdx ei = ee_interleaved(ctx->B, ctx->N);
        std::vector<double> r((size_t)ctx->D * N * ee_padded_bands(ctx->B));
        HIPCHK(hipMemcpy(r.data(), ctx->ee_final, sizeof(double) * r.size(), hipMemcpyDeviceToHost));
        for (int b = 0; b < ctx->B; b++)
            for (int d = 0; d < ctx->D; d++)
                for (size_t c = 0; c < N; c++) h[((size_t)b * ctx->D + d) * N + c] = r[ei.at(b, d, (int)c)];
    } else {
        HIPCHK(hipMemcpy(h.data(), ctx->ee_final, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
    }
    to_external(ctx, h.data(), ee, 1, S);
    return BTE_OK;
}

int bte_get_stats(bte_ctx* ctx, bte_stats* st) {
    if (!ctx || !st) return BTE_ERR_ARG;
    std::memset(st, 0, sizeof(*st));
    st->n_kernels = KA_COUNT;
    for (int k = 0; k < KA_COUNT; k++) {
        st->names[k] = kKernelNames[k];
        st->ms[k] = ctx->ms[k];
        st->calls[k] = ctx->calls[k];
    }
    st->spmv_rows = ctx->spmv_rows;
    st->ell_width = ctx->W;
    st->max_levels = ctx->max_levels;
    st->lagged_edges = ctx->lagged;
    st->block = ctx->block_used;
    st->variant = ctx->variant_used;
    st->band_group = ctx->group_used;
    st->intensity_layout = ctx->run_inter ? 2 : 1;
    return BTE_OK;
}

int bte_debug_assemble(bte_ctx* ctx, int band, int dir, double* diag, double* off) {
    if (!ctx || !diag || !off) return BTE_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    int rc = prepare(ctx, false);
    if (rc) return rc;
    if (band < 0 || band >= ctx->B || dir < 0 || dir >= ctx->D) SETERR(BTE_ERR_ARG, "band/dir out of range");
    const int N = ctx->N, F = ctx->F, W = ctx->W;
    std::vector<double> dg(N), val((size_t)W * N);
    std::vector<int> col((size_t)W * N);
    HIPCHK(hipMemcpy(dg.data(), ctx->dg.as<double>() + (size_t)dir * N, sizeof(double) * N, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(val.data(), ctx->ell_val.as<double>() + (size_t)dir * W * N, sizeof(double) * W * N,
                     hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(col.data(), ctx->ell_col.as<int>() + (size_t)dir * W * N, sizeof(int) * W * N,
                     hipMemcpyDeviceToHost));
    const double vt = ctx->h_v[band] * ctx->h_tau[band];
    const double w = (ctx->dim > 1) ? ctx->h_omega[dir] : 1.0;
    for (int c = 0; c < N; c++) {
        const size_t e = ext_cell(ctx, c);  // rows go out in the caller's cell order
        diag[e] = vt * dg[c] + w * ctx->h_vol[c];
        for (int f = 0; f < F; f++) off[e * F + f] = 0;
        for (int k = 0; k < W; k++) {
            const int j = col[(size_t)k * N + c];
            if (j < 0) continue;
            for (int f = 0; f < F; f++)
                if (ctx->h_nbr[(size_t)c * F + f] == j) {
                    off[e * F + f] = vt * val[(size_t)k * N + c];
                    break;
                }
        }
    }
    return BTE_OK;
}

// The debug entry points overwrite the solver's state buffers: a solve in progress cannot continue
// after them, and its intensities can no longer be read.
static int upload_state(bte_ctx* ctx, const double* ee_prev, const double* density) {
    ctx->ee_final = nullptr;
    ctx->run_active = false;
    ctx->tsrc_ok = false;  // a density uploaded as such has no source temperature
    const int N = ctx->N, B = ctx->B;
    const size_t S = (size_t)B * ctx->D, SN = S * N;
    std::vector<double> ee(SN);
    to_internal(ctx, ee_prev, ee.data(), 1, S);
    HIPCHK(hipMemcpy(ctx->eeB.p, ee.data(), sizeof(double) * SN, hipMemcpyHostToDevice));
    std::vector<double> dT((size_t)B * N);
    for (int i = 0; i < N; i++)
        for (int b = 0; b < B; b++) dT[(size_t)i * B + b] = density[(size_t)ext_cell(ctx, i) * B + b];
    HIPCHK(hipMemcpy(ctx->dens.p, dT.data(), sizeof(double) * dT.size(), hipMemcpyHostToDevice));
    return BTE_OK;
}

int bte_debug_rhs(bte_ctx* ctx, const double* ee_prev, const double* density, double* rhs) {
    if (!ctx || !ee_prev || !density || !rhs) return BTE_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    int rc = prepare(ctx, false);
    if (rc) return rc;
    if ((rc = ensure_work(ctx, BTE_PREC_FP64))) return rc;
    if ((rc = upload_state(ctx, ee_prev, density))) return rc;
    const size_t SN = (size_t)ctx->B * ctx->D * ctx->N;
    if (ctx->has_diffuse)
        hipLaunchKernelGGL(k_diffuse, dim3(diffuse_grid(ctx)), dim3(256), 0, ctx->stream,
                           geo_of(ctx), quad_of(ctx), bnd_of(ctx), ctx->B, ctx->eeB.as<double>(), ctx->ediff.as<double>());
    hipLaunchKernelGGL(k_rhs, dim3(grid_for((long long)SN, 256)), dim3(256), 0, ctx->stream, geo_of(ctx), quad_of(ctx),
                       bnd_of(ctx), bands_of(ctx), ctx->eeB.as<double>(), ctx->dens.as<double>(), ctx->wb.as<double>());
    HIPCHK(hipGetLastError());
    std::vector<double> h(SN);
    HIPCHK(hipMemcpyAsync(h.data(), ctx->wb.p, sizeof(double) * SN, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    to_external(ctx, h.data(), rhs, 1, SN / ctx->N);
    return BTE_OK;
}

int bte_debug_inner(bte_ctx* ctx, const bte_opts* opts, const double* ee_prev, const double* density, double* x,
                    int* iters) {
    if (!ctx || !ee_prev || !density || !x) return BTE_ERR_ARG;
    bte_opts o;
    if (opts)
        o = *opts;
    else
        bte_default_opts(&o);
    HIPCHK(hipSetDevice(ctx->device));
    int rc = prepare(ctx, false);
    if (rc) return rc;
    if (o.precision == BTE_PREC_MIXED && (rc = ensure_fp32(ctx))) return rc;
    if ((rc = ensure_work(ctx, o.precision))) return rc;
    if ((rc = upload_state(ctx, ee_prev, density))) return rc;
    const int S = ctx->B * ctx->D;
    const size_t SN = (size_t)S * ctx->N;
    HIPCHK(hipMemsetAsync(ctx->iters.p, 0, sizeof(int) * S, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->fails.p, 0, sizeof(int) * S, ctx->stream));
    if (ctx->has_diffuse)
        hipLaunchKernelGGL(k_diffuse, dim3(diffuse_grid(ctx)), dim3(256), 0, ctx->stream,
                           geo_of(ctx), quad_of(ctx), bnd_of(ctx), ctx->B, ctx->eeB.as<double>(), ctx->ediff.as<double>());
    Shape shape;
    if ((rc = pick_shape(ctx, &o, shape, false))) return rc;
    launch_solve(ctx, &o, shape, ctx->eeB.as<double>(), ctx->eeA.as<double>(), ctx->dens.as<double>());
    HIPCHK(hipGetLastError());
    std::vector<double> h(SN);
    HIPCHK(hipMemcpyAsync(h.data(), ctx->eeA.p, sizeof(double) * SN, hipMemcpyDeviceToHost, ctx->stream));
    if (iters) HIPCHK(hipMemcpyAsync(iters, ctx->iters.p, sizeof(int) * S, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    to_external(ctx, h.data(), x, 1, S);
    return BTE_OK;
}

// ---- uneven batch of problems ----------------------------------------------------------------

struct bte_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<bte_ctx*> ctx;
    DBuf desc, items, mptr, mout;
    DBuf idesc;                  // per iteration: IterDesc of each active problem, then step offsets
    std::vector<unsigned char> h_idesc;
    // per iteration, the problems' recorded mixing steps and the descriptors of their launches
    // (host copies kept until the next synchronisation of the batch stream)
    std::vector<AaRec> aa_rec;
    DBuf aa_first, aa_second;
    std::vector<unsigned char> h_aa_first, h_aa_second;
    std::vector<int> items_act;  // the active problems b->items and b->mptr were built for
    double* h_m = nullptr;  // pinned [K] margins
    int ring = 0, block = 256, n_items = 0, max_levels = 0;
    double seconds = 0;
    bool timing = false;               // bte_batch_solve with opts.timing: events around the launches
    double ms_rhs = 0, ms_sweep = 0;   // summed hipEvent milliseconds of the last solve
    long long n_rhs = 0, n_sweep = 0;
};

#define BCHK(expr)                                                                                \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            b->err = std::string(#expr) + ": " + hipGetErrorString(e_);                           \
            return BTE_ERR_HIP;                                                                   \
        }                                                                                         \
    } while (0)

int bte_batch_create(int n, bte_ctx* const* ctxs, bte_batch** out) {
    if (!out || n <= 0 || !ctxs) return BTE_ERR_ARG;
    *out = nullptr;
    for (int k = 0; k < n; k++)
        if (!ctxs[k] || ctxs[k]->device != ctxs[0]->device) return BTE_ERR_ARG;
    bte_batch* b = new bte_batch();
    b->device = ctxs[0]->device;
    b->ctx.assign(ctxs, ctxs + n);
    if (hipSetDevice(b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&b->h_m, sizeof(double) * n, hipHostMallocDefault) != hipSuccess) {
        if (b->stream) (void)hipStreamDestroy(b->stream);
        delete b;
        return BTE_ERR_HIP;
    }
    *out = b;
    return BTE_OK;
}

void bte_batch_destroy(bte_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->h_m) (void)hipHostFree(b->h_m);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char* bte_batch_last_error(const bte_batch* b) { return b ? b->err.c_str() : "null batch"; }

int bte_batch_get_stats(const bte_batch* b, bte_stats* st) {
    if (!b || !st) return BTE_ERR_ARG;
    std::memset(st, 0, sizeof(*st));
    st->n_kernels = KA_COUNT;
    for (int k = 0; k < KA_COUNT; k++) st->names[k] = kKernelNames[k];
    st->ms[KA_RHS] = b->ms_rhs;
    st->calls[KA_RHS] = b->n_rhs;
    st->ms[KA_SOLVE] = b->ms_sweep;
    st->calls[KA_SOLVE] = b->n_sweep;
    st->max_levels = b->max_levels;
    st->block = b->block;
    st->variant = b->ring > 0 ? 4 : 2;
    return BTE_OK;
}

// Runs the member contexts on the batch stream for the duration of a call.
struct BatchStreams {
    std::vector<bte_ctx*>& ctx;
    std::vector<hipStream_t> saved;
    BatchStreams(std::vector<bte_ctx*>& c, hipStream_t s) : ctx(c) {
        for (bte_ctx* x : ctx) {
            saved.push_back(x->stream);
            x->stream = s;
        }
    }
    ~BatchStreams() {
        for (size_t k = 0; k < ctx.size(); k++) ctx[k]->stream = saved[k];
    }
};

}  // extern "C"

template <int BLOCK>
static void launch_batch_b(bte_batch* b, int ring) {
    const SweepDesc* desc = b->desc.as<SweepDesc>();
    const int2* items = b->items.as<int2>();
    if (ring > 0) {
        const size_t lds = ring_lds_bytes(ring, b->max_levels);
        (void)hipFuncSetAttribute((const void*)k_sweep_batch<BLOCK, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
        hipLaunchKernelGGL((k_sweep_batch<BLOCK, true>), dim3(b->n_items), dim3(BLOCK), lds, b->stream, desc, items,
                           b->n_items, ring - 1);
    } else {
        hipLaunchKernelGGL((k_sweep_batch<BLOCK, false>), dim3(b->n_items), dim3(BLOCK), 0, b->stream, desc, items,
                           b->n_items, 0);
    }
}

extern "C" {

// One outer iteration of every active problem: boundary data, one batched sweep launch over all
// their systems, temperature update and margins; returns the margins (host) of the active problems.
static int batch_iteration(bte_batch* b, const std::vector<int>& act, std::vector<double>& m) {
    const int P = (int)act.size();
    std::vector<SweepDesc> hd(act.size());
    std::vector<std::pair<int, int2>> order;  // (levels, item)
    std::vector<const double*> mp(act.size());
    // the per-problem steps around the sweep, one launch per step for all problems: descriptors,
    // then the work-group offsets of the diffuse, interior rhs, boundary rhs, band temperature and
    // cell temperature steps (P + 1 each), uploaded in one copy
    enum { ST_DIFF, ST_RHS, ST_BND, ST_TBAND, ST_TCELL, ST_N };
    b->h_idesc.resize(sizeof(IterDesc) * P + sizeof(int) * ST_N * (P + 1));
    IterDesc* id = (IterDesc*)b->h_idesc.data();
    int* offs = (int*)(b->h_idesc.data() + sizeof(IterDesc) * P);
    for (int k = 0; k < ST_N; k++) offs[k * (P + 1)] = 0;
    for (size_t a = 0; a < act.size(); a++) {
        bte_ctx* ctx = b->ctx[act[a]];
        const int par = ctx->run_iter & 1;
        double* buf[2] = {ctx->eeA.as<double>(), ctx->eeB.as<double>()};
        IterDesc& e = id[a];
        e.g = geo_of(ctx);
        e.q = quad_of(ctx);
        e.bd = bnd_of(ctx);
        e.bands = bands_of(ctx);
        e.o = ops_of(ctx);
        e.ee_prev = buf[par];
        e.ee_curr = buf[1 - par];
        e.ediff = ctx->ediff.as<double>();
        e.rhs = ctx->wb.as<double>();
        e.bnd_list = ctx->bnd_list.as<int>();
        e.n_bnd_list = ctx->n_bnd_list;
        e.Tband = ctx->Tband.as<double>();
        e.T = ctx->T.as<double>();
        e.dens = ctx->dens.as<double>();
        e.tsrc = ctx->tsrc.as<double>();
        e.tsrc_rhs = ctx->tsrc_ok ? ctx->tsrc.as<double>() : nullptr;
        e.partial = ctx->partial.as<double>();
        e.margin = ctx->margin.as<double>();
        e.omega = ctx->omega.as<double>();
        e.nblk_tband = grid_for((long long)ctx->B * ctx->N, 256);
        e.scale = 1.0 / ctx->B / ctx->N;
        // work-groups per step, as the single-problem launches partition it
        const int nb[ST_N] = {ctx->has_diffuse ? diffuse_grid(ctx) : 0,
                              b->ring > 0 ? grid_for((long long)ctx->D * lo_stride(ctx->N), 256) : 0,
                              (b->ring > 0 && ctx->n_bnd_list > 0)
                                  ? rhs_bnd_grid(ctx) : 0,
                              e.nblk_tband, grid_for(ctx->N, 256)};
        for (int k = 0; k < ST_N; k++) offs[k * (P + 1) + a + 1] = offs[k * (P + 1) + a] + nb[k];
        SweepDesc& d = hd[a];  // per iteration: the intensity buffers alternate
        d.g = geo_of(ctx);
        d.q = quad_of(ctx);
        d.bd = bnd_of(ctx);
        d.bands = bands_of(ctx);
        d.o = ops_of(ctx);
        d.w = work_of(ctx, buf[1 - par]);
        d.ee_prev = buf[par];
        d.dens = ctx->dens.as<double>();
        mp[a] = ctx->margin.as<double>();
    }
    // the work items change only when a problem leaves the batch
    std::vector<int2> items;
    if (act != b->items_act) {
        // One group per (problem, direction): its bands share the direction's operator records.
        // Work-group t runs on XCD t % 8, so each XCD gets whole groups in consecutive slots of its
        // own (t = 8 j + x): a direction's bands sweep at the same time on one XCD and read the
        // records from its L2.  Groups go longest level chain first, each to the XCD with the
        // fewest items so far; the XCDs' shorter sequences are padded with empty items (x = -1).
        for (size_t a = 0; a < act.size(); a++) {
            const bte_ctx* ctx = b->ctx[act[a]];
            for (int d = 0; d < ctx->D; d++) order.push_back({ctx->h_lvl_cnt[d], make_int2((int)a, d)});
        }
        std::stable_sort(order.begin(), order.end(), [](const auto& x, const auto& y) { return x.first > y.first; });
        std::vector<int2> xq[kXcds];
        for (const auto& g : order) {
            int x = 0;
            for (int k = 1; k < kXcds; k++)
                if (xq[k].size() < xq[x].size()) x = k;
            const bte_ctx* ctx = b->ctx[act[g.second.x]];
            for (int bd = 0; bd < ctx->B; bd++) xq[x].push_back(make_int2(g.second.x, bd * ctx->D + g.second.y));
        }
        size_t len = 0;
        for (int x = 0; x < kXcds; x++) len = std::max(len, xq[x].size());
        items.assign(len * kXcds, make_int2(-1, 0));
        for (int x = 0; x < kXcds; x++)
            for (size_t j = 0; j < xq[x].size(); j++) items[j * kXcds + x] = xq[x][j];
        b->n_items = (int)items.size();
        BCHK(b->items.ensure(sizeof(int2) * items.size()));
        BCHK(b->mptr.ensure(sizeof(double*) * mp.size()));
        BCHK(hipMemcpyAsync(b->items.p, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice, b->stream));
        BCHK(hipMemcpyAsync(b->mptr.p, mp.data(), sizeof(double*) * mp.size(), hipMemcpyHostToDevice, b->stream));
        b->items_act = act;
    }
    BCHK(b->idesc.ensure(b->h_idesc.size()));
    BCHK(hipMemcpyAsync(b->idesc.p, b->h_idesc.data(), b->h_idesc.size(), hipMemcpyHostToDevice, b->stream));
    cmem<IterDesc>* did = (cmem<IterDesc>*)b->idesc.p;
    const int* doff = (const int*)((const char*)b->idesc.p + sizeof(IterDesc) * P);
    auto multi = [&](auto kernel, int step, int block) {
        const int total = offs[step * (P + 1) + P];
        if (total > 0)
            hipLaunchKernelGGL(kernel, dim3(total), dim3(block), 0, b->stream, did, doff + step * (P + 1), P);
    };
    BCHK(b->desc.ensure(sizeof(SweepDesc) * hd.size()));
    BCHK(b->mout.ensure(sizeof(double) * mp.size()));
    BCHK(hipMemcpyAsync(b->desc.p, hd.data(), sizeof(SweepDesc) * hd.size(), hipMemcpyHostToDevice, b->stream));
    hipEvent_t ev[3] = {};
    if (b->timing)
        for (auto& e : ev) BCHK(hipEventCreate(&e));
    if (b->timing) BCHK(hipEventRecord(ev[0], b->stream));
    // boundary data, then the right-hand sides in level order (interior, then boundary cells)
    multi(k_diffuse_multi, ST_DIFF, 256);
    multi(k_rhs_levels_multi, ST_RHS, 256);
    multi(k_rhs_bnd_multi, ST_BND, 64);
    if (b->timing) BCHK(hipEventRecord(ev[1], b->stream));
    switch (b->block) {
        case 64: launch_batch_b<64>(b, b->ring); break;
        case 128: launch_batch_b<128>(b, b->ring); break;
        case 512: launch_batch_b<512>(b, b->ring); break;
        default: launch_batch_b<256>(b, b->ring); break;
    }
    if (b->timing) {
        BCHK(hipEventRecord(ev[2], b->stream));
        BCHK(hipEventSynchronize(ev[2]));
        float t01 = 0, t12 = 0;
        BCHK(hipEventElapsedTime(&t01, ev[0], ev[1]));
        BCHK(hipEventElapsedTime(&t12, ev[1], ev[2]));
        for (auto& e : ev) (void)hipEventDestroy(e);
        b->ms_rhs += t01;
        b->ms_sweep += t12;
        b->n_rhs++;
        b->n_sweep++;
    }
    // band and cell temperatures, emission densities and margins
    multi(k_tband_multi, ST_TBAND, 256);
    multi(k_tcell_multi, ST_TCELL, 256);
    hipLaunchKernelGGL(k_margin_multi, dim3(P), dim3(256), 0, b->stream, did);
    hipLaunchKernelGGL(k_gather_scalars, dim3(1), dim3(256), 0, b->stream, (int)mp.size(),
                       b->mptr.as<const double*>(), b->mout.as<double>());
    BCHK(hipGetLastError());
    BCHK(hipMemcpyAsync(b->h_m, b->mout.p, sizeof(double) * mp.size(), hipMemcpyDeviceToHost, b->stream));
    BCHK(hipStreamSynchronize(b->stream));  // also keeps hd / items alive until the copies are done
    m.assign(b->h_m, b->h_m + mp.size());
    return BTE_OK;
}

}  // extern "C"

// Appends n records of T to a host staging buffer at an 8-byte aligned offset; returns the offset.
template <class T>
static size_t stage(std::vector<unsigned char>& h, const T* v, size_t n) {
    const size_t at = (h.size() + 7) & ~(size_t)7;
    h.resize(at + sizeof(T) * n);
    if (n) std::memcpy(h.data() + at, v, sizeof(T) * n);
    return at;
}

// The recorded first halves of the problems' mixing steps: the gathers, then the dot products and
// their reductions, one launch each for all problems.
static int batch_aa_launch(bte_batch* b) {
    std::vector<AaGather> g;
    std::vector<AaDots> d;
    std::vector<int> og{0}, od{0}, orr{0};
    for (const AaRec& r : b->aa_rec) {
        if (r.gather) {
            g.push_back(r.ga);
            og.push_back(og.back() + grid_for(r.ga.N + (long long)r.ga.S * r.ga.nb, 256));
        }
        if (r.dots) {
            d.push_back(r.da);
            od.push_back(od.back() + kAaDotBlocks * ((r.da.ncols + kAaDotG - 1) / kAaDotG + 1));
            orr.push_back(orr.back() + std::min(2 * r.da.ncols + 1, 128));
        }
    }
    if (g.empty()) return BTE_OK;
    std::vector<unsigned char>& h = b->h_aa_first;
    h.clear();
    const size_t at_g = stage(h, g.data(), g.size()), at_d = stage(h, d.data(), d.size());
    const size_t at_og = stage(h, og.data(), og.size()), at_od = stage(h, od.data(), od.size());
    const size_t at_or = stage(h, orr.data(), orr.size());
    BCHK(b->aa_first.ensure(h.size()));
    BCHK(hipMemcpyAsync(b->aa_first.p, h.data(), h.size(), hipMemcpyHostToDevice, b->stream));
    const char* base = (const char*)b->aa_first.p;
    hipLaunchKernelGGL(k_state_gather_diff_multi, dim3(og.back()), dim3(256), 0, b->stream,
                       (cmem<AaGather>*)(base + at_g), (const int*)(base + at_og), (int)g.size());
    if (!d.empty()) {
        hipLaunchKernelGGL((k_aa_dots_multi<256, kAaDotG>), dim3(od.back()), dim3(256), 0, b->stream,
                           (cmem<AaDots>*)(base + at_d), (const int*)(base + at_od), (int)d.size());
        hipLaunchKernelGGL(k_aa_reduce_multi, dim3(orr.back()), dim3(256), 0, b->stream, (cmem<AaDots>*)(base + at_d),
                           (const int*)(base + at_or), (int)d.size());
    }
    BCHK(hipGetLastError());
    return BTE_OK;
}

// The recorded updates, one launch for all problems; each problem's coefficients travel in the same
// upload as the descriptors.
static int batch_aa_update(bte_batch* b) {
    std::vector<AaUpdate> u;
    std::vector<double> gam;
    std::vector<size_t> gat;
    std::vector<int> ou{0};
    for (const AaRec& r : b->aa_rec)
        if (r.update) {
            u.push_back(r.ua);
            gat.push_back(gam.size());
            gam.insert(gam.end(), r.coef.g, r.coef.g + std::max(r.ua.ncols, 1));
            ou.push_back(ou.back() + grid_for(r.ua.N + (long long)r.ua.S * r.ua.nb, 256));
        }
    if (u.empty()) return BTE_OK;
    std::vector<unsigned char>& h = b->h_aa_second;
    h.clear();
    const size_t at_u = stage(h, u.data(), u.size()), at_c = stage(h, gam.data(), gam.size());
    const size_t at_o = stage(h, ou.data(), ou.size());
    BCHK(b->aa_second.ensure(h.size()));
    const char* base = (const char*)b->aa_second.p;
    AaUpdate* hu = (AaUpdate*)(h.data() + at_u);
    for (size_t k = 0; k < u.size(); k++) hu[k].gamma = (const double*)(base + at_c) + gat[k];
    BCHK(hipMemcpyAsync(b->aa_second.p, h.data(), h.size(), hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(k_aa_update_multi, dim3(ou.back()), dim3(256), 0, b->stream, (cmem<AaUpdate>*)(base + at_u),
                       (const int*)(base + at_o), (int)u.size());
    BCHK(hipGetLastError());
    return BTE_OK;
}

extern "C" {

int bte_batch_solve(bte_batch* b, const bte_opts* opts, bte_result* results) {
    if (!b) return BTE_ERR_ARG;
    BCHK(hipSetDevice(b->device));
    bte_opts o;
    if (opts)
        o = *opts;
    else
        bte_default_opts(&o);
    o.use_graph = 0;
    o.intensity_layout = 1;  // the batch's sweep and per-problem kernels index band-major intensities
    b->timing = o.timing != 0;
    b->ms_rhs = b->ms_sweep = 0;
    b->n_rhs = b->n_sweep = 0;
    o.timing = 0;
    const int K = (int)b->ctx.size();
    BatchStreams streams(b->ctx, b->stream);
    int ring = 1 << 30;
    for (int k = 0; k < K; k++) {
        bte_ctx* ctx = b->ctx[k];
        int rc = bte_begin(ctx, &o);
        if (rc) {
            b->err = "problem " + std::to_string(k) + ": " + ctx->err;
            return rc;
        }
        if (!direct_path(ctx, &o)) {
            b->err = "problem " + std::to_string(k) + ": the batch needs the FP64 direct-sweep form (precision FP64, "
                     "sweep preconditioner, direct_sweep, an acyclic upwind graph)";
            return BTE_ERR_ARG;
        }
        ring = std::min(ring, ring_size(ctx) > 0 ? ring_size(ctx) : 0);
    }
    // one LDS ring size for the launch: the largest any problem needs, if every problem can use one
    int rmax = 0;
    b->max_levels = 0;
    for (int k = 0; k < K; k++) {
        if (ring > 0) rmax = std::max(rmax, ring_size(b->ctx[k]));
        b->max_levels = std::max(b->max_levels, b->ctx[k]->max_levels);
    }
    b->ring = (ring > 0 && ring_lds_bytes(rmax, b->max_levels) <= (size_t)b->ctx[0]->max_lds) ? rmax : 0;
    b->block = (o.block == 64 || o.block == 128 || o.block == 512) ? o.block : 256;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<int> act;
    for (int k = 0; k < K; k++) act.push_back(k);
    b->items_act.clear();  // the members may have changed since the last call
    std::vector<double> m;
    while (!act.empty()) {
        int rc = batch_iteration(b, act, m);
        if (rc) return rc;
        std::vector<int> next;
        // the bookkeeping of every problem, with its mixing step recorded; then the steps of all of
        // them, one launch per kernel, with one synchronisation for their dot products
        b->aa_rec.assign(act.size(), AaRec{});
        auto fail = [&](int a) {
            b->err = "problem " + std::to_string(act[a]) + ": " + b->ctx[act[a]]->err;
            return rc;
        };
        std::vector<int> fin, plain;
        for (size_t a = 0; a < act.size(); a++) {
            bte_ctx* ctx = b->ctx[act[a]];
            const int par = ctx->run_iter & 1;
            ctx->ee_final = par ? ctx->eeA.as<double>() : ctx->eeB.as<double>();
            ctx->run_iter++;
            bool counted = false, mix = false, finish = false;
            rc = after_judge(ctx, m[a], counted, mix);
            if (!rc && mix) rc = aa_launch(ctx, &finish, &b->aa_rec[a]);
            if (rc) return fail((int)a);
            if (finish)
                fin.push_back((int)a);
            else if (mix)
                plain.push_back((int)a);
        }
        if ((rc = batch_aa_launch(b))) return rc;
        for (int a : plain)
            if ((rc = after_mix(b->ctx[act[a]]))) return fail(a);
        if (!fin.empty()) {
            BCHK(hipStreamSynchronize(b->stream));
            for (int a : fin)
                if ((rc = aa_finish(b->ctx[act[a]], &b->aa_rec[a]))) return fail(a);
            if ((rc = batch_aa_update(b))) return rc;
            for (int a : fin)
                if ((rc = after_mix(b->ctx[act[a]]))) return fail(a);
        }
        for (size_t a = 0; a < act.size(); a++) {
            bte_ctx* ctx = b->ctx[act[a]];
            if (!ctx->run_converged && (int)ctx->run_margins.size() < o.max_iter) next.push_back(act[a]);
        }
        act.swap(next);
    }
    b->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    int status = BTE_OK;
    for (int k = 0; k < K; k++) {
        bte_ctx* ctx = b->ctx[k];
        ctx->run_solve_seconds = b->seconds;
        if (!results) continue;
        const int rc = bte_finish(ctx, &results[k]);
        if (rc && rc != BTE_ERR_NOCONV) {
            b->err = "problem " + std::to_string(k) + ": " + ctx->err;
            return rc;
        }
        if (rc) status = rc;
    }
    return status;
}

}  // extern "C"
