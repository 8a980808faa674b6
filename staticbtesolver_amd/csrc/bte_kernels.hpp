// bte_kernels.hpp -- gfx950 device code of the source iteration.
//
// Layout in HBM (one context, N cells, F faces per cell, D directions, B bands, S = B*D systems):
//   geometry     vol[N], area[N*F], normal[N*F*3], nbr[N*F]           (read by every system)
//   assembly     per direction, shared by all bands of that direction:
//                dg[D*N]            sum of A_f (S_d.n_f) over outgoing faces
//                ell_col/ell_val[(d*W+k)*N + c]  incoming interior faces, k < W = F-1 slots
//   sweep        per direction, cells in downwind level order:
//                lvl_ptr[lvl_off[d] .. + L_d], perm[d*N+i], swp_col/swp_val[(d*W+k)*N+i], swp_dg[d*N+i]
//   intensities  ee at b*sb + d*sd + c*sc (EeIdx): band-major ee[b][d][c] (the reference's
//                ee_curr[b]->data layout), or direction-major with a cell's bands interleaved,
//                ee[d][c][b] (the pipelined sweep in 3D: the co-scheduled bands of a direction then
//                write each cell's intensities together)
//   Krylov work  one slice of N per system and vector, in band-major system order
//
// Matrix of system (b, d), StaticBTESolver.cpp:122-160:
//   diag(c) = v_b tau_b dg[d][c] + w_d V_c        (w_d = omega_d in 2D/3D, 1 in 1D)
//   off(c, k) = v_b tau_b ell_val[d][k][c]       (column ell_col[d][k][c])
// The band enters only through the scalar v_b tau_b, so the stored operator is shared by bands.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bte {

constexpr int kMaxBC = 32;
constexpr double kPi = 3.14159265358979323846;

// Dot product without FMA contraction, summed x+y+z left to right: the upwind test d.n >= 0 must
// classify every face identically in every kernel (and on the host).
__device__ __forceinline__ double dot_nc(const double* a, const double* b) {
#pragma clang fp contract(off)
    double s = a[0] * b[0];
    s = s + a[1] * b[1];
    s = s + a[2] * b[2];
    return s;
}

struct Geo {
    int dim, N, F;
    const double* vol;
    const double* area;
    const double* normal;
    const int* nbr;
};

// Index of intensity (band b, direction d, cell c) in an intensity array: b sb + d sd + c sc.
// Band-major: sb = D N, sd = N, sc = 1.  Direction-major, bands interleaved: sb = 1, sd = N Bp,
// sc = Bp, with Bp = B rounded up to a multiple of 4, so that a cell's bands fill whole 32-byte
// sectors.
struct EeIdx {
    long long sb, sd, sc;
    __host__ __device__ size_t at(int b, int d, int c) const {
        return (size_t)(b * sb + d * sd + (long long)c * sc);
    }
};

__host__ __device__ inline EeIdx ee_band_major(int D, int N) { return EeIdx{(long long)D * N, N, 1}; }
__host__ __device__ inline int ee_padded_bands(int B) { return (B + 3) & ~3; }
__host__ __device__ inline EeIdx ee_interleaved(int B, int N) {
    return EeIdx{1, (long long)N * ee_padded_bands(B), ee_padded_bands(B)};
}

struct Quad {
    int D;
    const double* omega;
    const double* dir;
    const double* S;
    EeIdx ee;  // layout of the intensity arrays the kernels read
};

// ----------------------------------------------------------------------------------------------
// Assembly (StaticBTESolver.cpp:122-160), one thread per (direction, cell).  Writes the band-free
// parts of the operator.  *overflow counts rows with more than W incoming interior faces.
// ----------------------------------------------------------------------------------------------
__global__ void k_assemble(Geo g, Quad q, int W, double* __restrict__ dg, int* __restrict__ ell_col,
                           double* __restrict__ ell_val, int* __restrict__ overflow) {
    const long long total = (long long)q.D * g.N;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int d = (int)(t / g.N), c = (int)(t % g.N);
        const double* dv = q.dir + 3 * d;
        const double* sv = (g.dim > 1) ? q.S + 3 * d : dv;
        double out = 0;
        int k = 0;
        for (int f = 0; f < g.F; f++) {
            const double* n = g.normal + 3 * ((size_t)c * g.F + f);
            const double coef = g.area[(size_t)c * g.F + f] * dot_nc(sv, n);
            if (dot_nc(dv, n) >= 0) {
                out += coef;
            } else {
                const int nb = g.nbr[(size_t)c * g.F + f];
                if (nb >= 0) {
                    if (k < W) {
                        ell_col[((size_t)d * W + k) * g.N + c] = nb;
                        ell_val[((size_t)d * W + k) * g.N + c] = coef;
                    } else {
                        atomicAdd(overflow, 1);
                    }
                    k++;
                }
            }
        }
        for (; k < W; k++) {
            ell_col[((size_t)d * W + k) * g.N + c] = -1;
            ell_val[((size_t)d * W + k) * g.N + c] = 0;
        }
        dg[(size_t)d * g.N + c] = out;
    }
}

// Gathers the operator into downwind level order (sweep schedule built on the host from ell_col).
template <typename R>
__global__ void k_gather_sweep(int N, int D, int W, const int* __restrict__ perm, const double* __restrict__ dg,
                               const double* __restrict__ ell_val, R* __restrict__ swp_dg,
                               R* __restrict__ swp_val) {
    const long long total = (long long)D * N;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int d = (int)(t / N), i = (int)(t % N);
        const int c = perm[(size_t)d * N + i];
        swp_dg[(size_t)d * N + i] = (R)dg[(size_t)d * N + c];
        for (int k = 0; k < W; k++)
            swp_val[((size_t)d * W + k) * N + i] = (R)ell_val[((size_t)d * W + k) * N + c];
    }
}

// Packed level-order operand records of the pipelined sweeps, one 16-byte and one 32-byte record
// per (direction, position), so that a cell's operands are three 16-byte loads:
//   pki = {cell, upwind positions times 8 (byte offsets of doubles; an absent slot has INT_MAX,
//          which is never older than a ring and whose ring slot, the last, is read with
//          coefficient 0)},
//   pkd = {coefficients (absent: 0), then the diagonal face sum and the volume weight as far as
//          they fit: W = 1 {a0, dg, vw, 0}, W = 2 {a0, a1, dg, vw}, W = 3 {a0, a1, a2, dg}, the
//          volume weight then staying in swp_vw}.
__global__ void k_pack_sweep(int N, int D, int W, const int* __restrict__ perm, const int* __restrict__ spos,
                             const double* __restrict__ sval, const double* __restrict__ sdg,
                             const double* __restrict__ svw, int* __restrict__ pki, double* __restrict__ pkd) {
    const long long total = (long long)D * N;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int d = (int)(t / N), i = (int)(t % N);
        double a[3] = {0, 0, 0};
        pki[t * 4] = perm[t];
        for (int k = 0; k < 3; k++) {
            const int q = k < W ? spos[((size_t)d * W + k) * N + i] : -1;
            pki[t * 4 + 1 + k] = q < 0 ? 0x7fffffff : q * 8;
            if (q >= 0) a[k] = sval[((size_t)d * W + k) * N + i];
        }
        double* f = pkd + t * 4;
        f[0] = a[0];
        if (W == 1) {
            f[1] = sdg[t];
            f[2] = svw[t];
            f[3] = 0.0;
        } else if (W == 2) {
            f[1] = a[1];
            f[2] = sdg[t];
            f[3] = svw[t];
        } else {
            f[1] = a[1];
            f[2] = a[2];
            f[3] = sdg[t];
        }
    }
}

// Per-direction, level-ordered volume weights w_d V_c and a flag for cells with an incoming
// boundary face in direction d (the only cells whose right-hand side needs the face loop).
__global__ void k_sweep_aux(Geo g, Quad q, const int* __restrict__ perm, const double* __restrict__ w_eff,
                            double* __restrict__ swp_vw, unsigned char* __restrict__ swp_bnd) {
    const long long total = (long long)q.D * g.N;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int d = (int)(t / g.N), i = (int)(t % g.N);
        const int c = perm[(size_t)d * g.N + i];
        swp_vw[t] = w_eff[d] * g.vol[c];
        unsigned char bnd = 0;
        for (int f = 0; f < g.F; f++)
            if (g.nbr[(size_t)c * g.F + f] < 0 && dot_nc(q.dir + 3 * d, g.normal + 3 * ((size_t)c * g.F + f)) < 0)
                bnd = 1;
        swp_bnd[t] = bnd;
    }
}

template <typename R>
__global__ void k_convert(long long n, const double* __restrict__ in, R* __restrict__ out) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        out[t] = (R)in[t];
}

// ----------------------------------------------------------------------------------------------
// Boundary data
// ----------------------------------------------------------------------------------------------
struct Bnd {
    int nbc;
    int bci[kMaxBC];  // index = -k-1
    int bct[kMaxBC];  // type
    double bcT[kMaxBC];
    int n_bf;                // boundary faces (nbr < 0)
    const int* bf_cell;      // [n_bf]
    const int* bf_face;      // [n_bf]
    const int* bfid;         // [N*F] -> boundary face id or -1
    const int* mirror;       // [3*D] mirror direction for types 31, 32, 33 (-1 unsupported)
    const double* ediff;     // [B*n_bf] diffuse re-emission of the previous iterate
    int has_diffuse;
};

struct Bands {
    int B;
    const double* v;
    const double* tau;
    const double* C;
    const double* Lr;
    double solid_angle, T_ref;
};

// Diffuse (type 2) re-emitted intensity of each boundary face, StaticBTESolver.cpp:180-204.  It
// does not depend on the incoming direction, so it is formed once per (band, face) and iteration.
// With the bands of a cell interleaved in the intensity arrays (ee.sb == 1): one thread per (face,
// group of 4 bands), the groups of a face in neighbouring lanes, so that a group's four
// intensities of a direction are one 32-byte sector (the face's groups together read the cell's
// whole record) and the outgoing test and the weights are formed once for the group.  Band-major:
// one thread per (band, face), faces in neighbouring lanes (more threads for the scattered loads).
// Each band sums its outgoing terms in direction order either way.
// Body over work-group blk of nblk (blockDim.x threads each); k_diffuse runs it over the grid.
template <int G, class BndT>
__device__ __forceinline__ void diffuse_range_g(const Geo& g, const Quad& q, const BndT& bd, int B,
                                                const double* __restrict__ ee_prev, double* __restrict__ ediff,
                                                int blk, int nblk) {
    const int ng = (B + G - 1) / G;
    const int total = ng * bd.n_bf;
    for (int t = blk * blockDim.x + threadIdx.x; t < total; t += nblk * blockDim.x) {
        const int bf = G == 1 ? t % bd.n_bf : t / ng;
        const int b0 = G == 1 ? t / bd.n_bf : (t % ng) * G, nbg = min(G, B - b0);
        const int c = bd.bf_cell[bf], f = bd.bf_face[bf];
        const int nb = g.nbr[(size_t)c * g.F + f];
        bool diffuse = false;
        for (int k = 0; k < bd.nbc; k++) diffuse |= (bd.bci[k] == nb && bd.bct[k] == 2);
        if (!diffuse) {
            for (int j = 0; j < nbg; j++) ediff[(size_t)(b0 + j) * bd.n_bf + bf] = 0;
            continue;
        }
        const double* n = g.normal + 3 * ((size_t)c * g.F + f);
        double es[G] = {}, tw = 0;
        // the intensities of 8 directions are loaded together (a load under the outgoing test
        // would make each wait for the last); the sums take the outgoing terms in direction order
        constexpr int K = 8;
        for (int d0 = 0; d0 < q.D; d0 += K) {
            double ev[K][G];
#pragma unroll
            for (int k = 0; k < K; k++)
#pragma unroll
                for (int j = 0; j < G; j++)
                    ev[k][j] = (d0 + k < q.D && j < nbg) ? ee_prev[q.ee.at(b0 + j, d0 + k, c)] : 0.0;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int d2 = d0 + k;
                if (d2 >= q.D) break;
                const double dn = dot_nc(q.dir + 3 * d2, n);
                if (dn > 0) {
                    if (g.dim == 3) {
                        const double sn = dot_nc(q.S + 3 * d2, n);
#pragma unroll
                        for (int j = 0; j < G; j++) es[j] += ev[k][j] * sn * q.omega[d2];
                        tw += sn * q.omega[d2];
                    } else if (g.dim == 2) {
                        const double sn = dot_nc(q.S + 3 * d2, n);
#pragma unroll
                        for (int j = 0; j < G; j++) es[j] += ev[k][j] * sn;
                        tw += sn;
                    } else {
#pragma unroll
                        for (int j = 0; j < G; j++) es[j] += ev[k][j] * dn * q.omega[d2];
                    }
                }
            }
        }
        for (int j = 0; j < nbg; j++)
            ediff[(size_t)(b0 + j) * bd.n_bf + bf] = (g.dim != 1) ? es[j] / tw : es[j] / kPi;
    }
}

__host__ __device__ inline int diffuse_group(const EeIdx& ei) { return ei.sb == 1 ? 4 : 1; }

template <class BndT>
__device__ __forceinline__ void diffuse_range(const Geo& g, const Quad& q, const BndT& bd, int B,
                                              const double* __restrict__ ee_prev, double* __restrict__ ediff, int blk,
                                              int nblk) {
    if (diffuse_group(q.ee) == 4)
        diffuse_range_g<4>(g, q, bd, B, ee_prev, ediff, blk, nblk);
    else
        diffuse_range_g<1>(g, q, bd, B, ee_prev, ediff, blk, nblk);
}

__global__ void k_diffuse(Geo g, Quad q, Bnd bd, int B, const double* __restrict__ ee_prev,
                          double* __restrict__ ediff) {
    diffuse_range(g, q, bd, B, ee_prev, ediff, blockIdx.x, gridDim.x);
}

// Right-hand side of system (b, d) at cell c, StaticBTESolver.cpp:162-269.
template <class BndT>
__device__ __forceinline__ double rhs_at(const Geo& g, const Quad& q, const BndT& bd, const Bands& bands, int b, int d,
                                         int c, const double* __restrict__ ee_prev, const double* __restrict__ dens) {
    double re = 0;
    const double* dv = q.dir + 3 * d;
    const double* sv = (g.dim > 1) ? q.S + 3 * d : dv;
    const double vt = bands.v[b] * bands.tau[b];
    for (int f = 0; f < g.F; f++) {
        const int nb = g.nbr[(size_t)c * g.F + f];
        if (nb >= 0) continue;  // boundary indices are negative; interior faces carry no BC
        const double* n = g.normal + 3 * ((size_t)c * g.F + f);
        if (dot_nc(dv, n) >= 0) continue;
        const double a = vt * g.area[(size_t)c * g.F + f] * dot_nc(sv, n);
        for (int k = 0; k < bd.nbc; k++) {
            if (bd.bci[k] != nb) continue;
            const int type = bd.bct[k];
            if (type == 1) {
                re -= bands.C[b] / bands.solid_angle * (bd.bcT[k] - bands.T_ref) * a;
            } else if (type == 2) {
                re -= bd.ediff[(size_t)b * bd.n_bf + bd.bfid[(size_t)c * g.F + f]] * a;
            } else {
                const int slot = type - 31;  // 31, 32, 33 validated on the host
                const int m = bd.mirror[slot * q.D + d];
                re -= ee_prev[q.ee.at(b, m, c)] * a;
            }
        }
    }
    const double w = (g.dim != 1) ? q.omega[d] : 1.0;
    re += dens[(size_t)c * bands.B + b] * g.vol[c] * w;
    return re;
}

// Standalone RHS of every system (parity tests).
__global__ void k_rhs(Geo g, Quad q, Bnd bd, Bands bands, const double* __restrict__ ee_prev,
                      const double* __restrict__ dens, double* __restrict__ out) {
    const long long total = (long long)bands.B * q.D * g.N;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(t % g.N);
        const int s = (int)(t / g.N);
        out[t] = rhs_at(g, q, bd, bands, s / q.D, s % q.D, c, ee_prev, dens);
    }
}

// ----------------------------------------------------------------------------------------------
// Work-group reductions: wave64 butterfly, then the per-wave sums are read back from LDS by every
// thread in the same order, so all threads hold the bit-identical total (uniform control flow).
// ----------------------------------------------------------------------------------------------
template <int BLOCK, int K>
__device__ __forceinline__ void block_sum(double* v, double* lds) {
    constexpr int NW = BLOCK / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double x = v[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        v[k] = x;
    }
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) lds[k * NW + wid] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) s += lds[k * NW + w];
        v[k] = s;
    }
}

// ----------------------------------------------------------------------------------------------
// Batched inner solve: one work-group per system, the whole BiCGSTAB loop inside the kernel
// (per-system convergence, no host round trips).  Systems are s = b*D + d.
// ----------------------------------------------------------------------------------------------
// Row stride of the level-order right-hand sides rhs_lo (cells, N rounded up to 64): every wave
// of the pass that writes them stores whole, aligned 512-byte segments (a partial line written to
// HBM costs a read of it as well).
__host__ __device__ inline int lo_stride(int N) { return (N + 63) & ~63; }

struct Ops {
    int N, D, W;
    int Np;               // lo_stride(N)
    const double* vol;
    const double* omega;  // w_d (already 1 in 1D)
    const double* vt;     // [B] v_b * tau_b
    const double* band_C;  // [B] C_b: emission density C_b / solid_angle * (T - T_ref)
    double solid_angle, T_ref;
    const double* dg;     // [D*N]
    const int* ell_col;   // [(d*W+k)*N + c]
    const double* ell_val;
    const float* ell_val32;   // FP32 copy for the mixed-precision inner solve
    const double* dg32src;    // unused placeholder for symmetry
    const float* dg32;        // [D*N] FP32 copy
    // sweep schedule
    const int* lvl_off;   // [D] offset of direction d's level pointer list
    const int* lvl_cnt;   // [D] number of levels of direction d
    const int* lvl_ptr;   // concatenated [sum(L_d + 1)]
    const int* perm;      // [D*N]
    const int* swp_col;   // [(d*W+k)*N + i]
    const double* swp_val;
    const double* swp_dg;
    const float* swp_val32;
    const float* swp_dg32;
    const double* swp_vw;          // [D*N] w_d V_c in level order
    const unsigned char* swp_bnd;  // [D*N] 1 if the cell has an incoming boundary face
    const int* swp_pos;   // [(d*W+k)*N + i] level-order position of the upwind neighbour (-1: none)
    int ring_sync;        // band-batched pipelined sweeps: HBM writes made visible every ring_sync levels
    // pipelined sweeps with a ring of ring_r values: lvl_sync[lvl_off[d] + l] bit 1 when the sweep
    // of direction d must make its HBM writes visible after level l (a later level reads, from HBM, a
    // value older than the ring written at or before l), bit 0 when level l reads such values;
    // dir_fb[d] bit 0 when direction d has such reads at all.  Bits 3, 2 and dir_fb bit 1: the same
    // for the wide ring of ring_r2 values (0: none), which a launch selects with ring_sel = 1.
    const unsigned char* lvl_sync;
    const int* dir_fb;
    int ring_r;
    int ring_r2;
    int ring_sel;
    // smallest ring the selected flags hold for
    __host__ __device__ int flag_ring() const { return ring_sel ? ring_r2 : ring_r; }
    // packed level-order records of the pipelined sweeps (k_pack_sweep), 4 entries per (d, i)
    const int* swp_pki;
    const double* swp_pkd;
    // layout of the intensity arrays (w.x): band-major for every solver form; the pipelined sweep
    // (variant 4, sweep_ring_system) also writes the interleaved one
    EeIdx ee;
};

struct Work {
    double* x;     // ee_curr [S*N]
    double* b;     // rhs [S*N]
    const double* tlo;  // [D][Np] source temperature of the cell at each level-order position
    double* r64;   // FP64 residual (mixed precision)
    void* r;       // Krylov vectors of type R [S*N] each
    void* rh;
    void* p;
    void* v;
    void* ph;
    void* sh;
    void* t;
    void* y;       // mixed precision: correction of the current refinement step
    void* f;       // mixed precision: scaled FP64 residual rounded to FP32 (inner right-hand side)
    int* iters;    // [S] accumulated inner iterations
    int* fails;    // [S] accumulated iteration-limit hits
};

struct SolveCfg {
    int S;
    int maxit;
    double tol;
    int refine_steps;
    int reference_semantics;  // Jacobi path: reproduce bicgstab.hpp's first step and restarts
    double tol32;             // mixed precision: relative residual target of each FP32 inner solve
};

template <typename R>
struct SysView {
    int N, W, d;
    R vt;           // v_b tau_b
    R wv;           // w_d
    const double* vol;
    const R* dg;    // dg[d*N ..]
    const int* col; // ell_col + d*W*N
    const R* val;
    // sweep
    int L;
    const int* lp;
    const int* perm;
    const int* scol;
    const R* sval;
    const R* sdg;
};

template <typename R>
__device__ __forceinline__ R diag_of(const SysView<R>& sv, R dgc, int c) {
    return sv.vt * dgc + sv.wv * (R)sv.vol[c];
}

// x / d for the direct sweeps' diagonal solve: the hardware reciprocal estimate, one Newton step
// r = r0 + r0 (1 - d r0), and one residual correction q + r (x - d q) with q = x r (Markstein's
// quotient correction, both fused). d is the diagonal vt*dgc + w*vol of a cell, a positive normal
// number, so the operand scaling and special-case fix-up of the IEEE division sequence are not
// needed. Six operations instead of eleven; on a million random pairs with d in [1e-16, 1e3] it
// returned the correctly rounded quotient every time. Every direct-sweep variant uses it, so
// their results stay identical to each other.
__device__ __forceinline__ double sweep_div(double x, double d) {
    const double r0 = __builtin_amdgcn_rcp(d);
    const double r = __builtin_fma(r0, __builtin_fma(-d, r0, 1.0), r0);
    const double q = x * r;
    return __builtin_fma(r, __builtin_fma(-d, q, x), q);
}

// Streaming loops of the Krylov solver.  Thread t visits cells t, t + BLOCK, t + 2 BLOCK, ... in
// that order (so per-thread partial sums are accumulated in a fixed order); full groups of U cells
// are handed to the body at once, which issues all their loads before any arithmetic: U times the
// bytes in flight per wave for the same occupancy.
template <int K>
struct Unroll {
    static constexpr int value = K;
};
template <bool B>
struct BoolC {
    static constexpr bool value = B;
};

template <int BLOCK, int U, class F>
__device__ __forceinline__ void for_cells(int N, F&& f) {
    int c = threadIdx.x;
    for (; c + (U - 1) * BLOCK < N; c += U * BLOCK) f(Unroll<U>{}, c);
    for (; c < N; c += BLOCK) f(Unroll<1>{}, c);
}

// Off-diagonal part of (A x) for K cells c0 + k BLOCK: off[k] = sum_f a_f x_{j_f} (without the
// v_b tau_b factor).  Operator slots beyond the system's width and missing upwind neighbours
// (column -1) contribute nothing; the gather of a missing neighbour reads the cell itself, so no
// load is conditional.
template <typename R, int BLOCK, int K>
__device__ __forceinline__ void offdiag_cells(const SysView<R>& sv, const R* x, int c0, R (&off)[K]) {
    const int N = sv.N, W = sv.W;
    int j[K][3];
    R a[K][3];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int c = c0 + k * BLOCK;
#pragma unroll
        for (int f = 0; f < 3; f++) {
            j[k][f] = (f < W) ? sv.col[(size_t)f * N + c] : -1;
            a[k][f] = (f < W) ? sv.val[(size_t)f * N + c] : (R)0;
        }
    }
    R g[K][3];
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int f = 0; f < 3; f++) g[k][f] = x[j[k][f] >= 0 ? j[k][f] : c0 + k * BLOCK];
#pragma unroll
    for (int k = 0; k < K; k++) {
        R o = 0;
#pragma unroll
        for (int f = 0; f < 3; f++)
            if (j[k][f] >= 0) o += a[k][f] * g[k][f];
        off[k] = o;
    }
}

// y = A x for K cells c0 + k BLOCK; acc(k, c, y_c) sees each result in order.  With the Jacobi
// preconditioner x = z / diag, and the diagonal term diag_c x_c is taken as z_c (dz = z: one
// stream instead of the diagonal's two and x_c; equal to diag_c x_c up to the rounding of the
// division); otherwise (DZ false) it is formed from the operator.
template <typename R, int BLOCK, int K, bool DZ, class Acc>
__device__ __forceinline__ void spmv_cells(const SysView<R>& sv, const R* x, const R* dz, R* __restrict__ y, int c0,
                                           Acc&& acc) {
    R dt[K];
    if constexpr (DZ) {
#pragma unroll
        for (int k = 0; k < K; k++) dt[k] = dz[c0 + k * BLOCK];
    } else {
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int c = c0 + k * BLOCK;
            dt[k] = (sv.vt * sv.dg[c] + sv.wv * (R)sv.vol[c]) * x[c];
        }
    }
    R off[K];
    offdiag_cells<R, BLOCK, K>(sv, x, c0, off);
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int c = c0 + k * BLOCK;
        const R yc = dt[k] + sv.vt * off[k];
        y[c] = yc;
        acc(k, c, yc);
    }
}

// (A x)[c]
template <typename R>
__device__ __forceinline__ R spmv_row(const SysView<R>& sv, const R* x, int c) {
    R acc = diag_of(sv, sv.dg[c], c) * x[c];
    R off = 0;
#pragma unroll 3
    for (int k = 0; k < sv.W; k++) {
        const int j = sv.col[(size_t)k * sv.N + c];
        if (j >= 0) off += sv.val[(size_t)k * sv.N + c] * x[j];
    }
    return acc + sv.vt * off;
}

// y = M^-1 v: forward substitution in downwind level order (exact inverse when no edge is lagged).
template <typename R, int BLOCK>
__device__ __forceinline__ void sweep(const SysView<R>& sv, const R* v, R* y) {
    for (int l = 0; l < sv.L; l++) {
        const int beg = sv.lp[l], end = sv.lp[l + 1];
        for (int i = beg + (int)threadIdx.x; i < end; i += BLOCK) {
            const int c = sv.perm[i];
            R s = 0;
#pragma unroll 3
            for (int k = 0; k < sv.W; k++) {
                const int j = sv.scol[(size_t)k * sv.N + i];
                if (j >= 0) s += sv.sval[(size_t)k * sv.N + i] * y[j];
            }
            y[c] = (v[c] - sv.vt * s) / diag_of(sv, sv.sdg[i], c);
        }
        __syncthreads();
    }
}

// BiCGSTAB on A x = rhs for one system, right-preconditioned.  On entry r holds rhs; x is
// overwritten.  Two variants:
//  - reference semantics (Jacobi path): bicgstab.hpp:146-298 -- p starts as r, q = 0, the first
//    step forms p = (1 + rho) r; breakdowns (|rho| or |<r_hat, q>| < 1e-200) restart from x = 0
//    with r_hat = 6.8 rhs + 0.6; stop when ||r|| / ||r0|| < tol on the recursive residual.
//  - standard form (sweep path): p = 0, v = 0 so the first step is p = r; an early exit when
//    ||s|| / ||r0|| < tol (with the exact sweep this ends after one sweep and one product).
// Returns the iterations used (maxit + 1 if the limit was hit without convergence, or if
// stall_window > 0 and the residual norm set no new minimum for that many iterations).
//
// Five passes over the vectors per iteration, each fusing what needs no global reduction in
// between: (1) p update, with the Jacobi preconditioner applied in the same pass; (2) v = A p^
// with <r_hat, v>; (3) s = r - alpha v with ||s||^2, and s^ = s / diag (Jacobi); (4) t = A s^ with
// <t, s> and <t, t>; (5) the x and r updates with ||r||^2 and <r_hat, r>.  The arithmetic of every
// element and the order of every per-thread sum are those of the unfused sequence.
template <typename R, int BLOCK, bool SWEEP>
__device__ int bicgstab(const SysView<R>& sv, const R* rhs, R* x, R* r, R* rh, R* p, R* v, R* ph, R* sh, R* t,
                        int maxit, double tol, bool refsem, double* lds, int stall_window = 0) {
    // cells per group in the streaming passes and in the operator products: FP32 groups twice as
    // many cells for the same bytes in flight; 1024-wide work-groups get half (their 128-register
    // budget would otherwise spill)
    constexpr int US = ((BLOCK >= 1024) ? 2 : 4) * (sizeof(R) == 4 ? 2 : 1);
    constexpr int UM = ((BLOCK >= 1024) ? 1 : 2) * (sizeof(R) == 4 ? 2 : 1);
    const int N = sv.N;
    double red[2];
    // setup: x = 0, r = rhs, r_hat = r, p = r (reference) or 0, v = 0
    red[0] = 0;
    for_cells<BLOCK, US>(N, [&](auto kc, int c0) {
        constexpr int K = decltype(kc)::value;
        R bc[K];
#pragma unroll
        for (int k = 0; k < K; k++) bc[k] = rhs[c0 + k * BLOCK];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int c = c0 + k * BLOCK;
            x[c] = 0;
            r[c] = bc[k];
            rh[c] = bc[k];
            p[c] = refsem ? bc[k] : (R)0;
            v[c] = 0;
            red[0] += (double)bc[k] * (double)bc[k];
        }
    });
    block_sum<BLOCK, 1>(red, lds);
    const double nrm0 = sqrt(red[0]);
    double nrm = nrm0;
    double rho = 1, rhop = 1, alpha = 1, omega = 1;
    double rho_next = red[0];  // <r_hat, r> with r_hat = r
    int it = 0;
    bool converged = false;
    double best = nrm0;
    int best_it = 0;
    for (int i = 1; i <= maxit; i++) {
        it = i;
        if (nrm < 1e-50) {
            converged = true;
            break;
        }
        rhop = rho;
        rho = rho_next;
        if (rho < 1e-200 && rho > -1e-200) {
            // restart (bicgstab.hpp:170-180)
            rho = alpha = omega = 1;
            for (int c = threadIdx.x; c < N; c += BLOCK) {
                const R bc = rhs[c];
                x[c] = 0;
                v[c] = 0;
                p[c] = 0;
                r[c] = bc;
                rh[c] = refsem ? (R)(bc * (R)6.8 + (R)0.6) : bc;
            }
            __syncthreads();
            red[0] = 0;
            for (int c = threadIdx.x; c < N; c += BLOCK) red[0] += (double)rh[c] * (double)r[c];
            block_sum<BLOCK, 1>(red, lds);
            rho_next = red[0];
            continue;
        }
        const double beta = (rho / rhop) * (alpha / omega);
        // (1) p = (p - omega v) * beta + r; Jacobi: p^ = p / diag
        for_cells<BLOCK, US>(N, [&](auto kc, int c0) {
            constexpr int K = decltype(kc)::value;
            R pv[K], vv[K], rv[K], dgv[K];
            double vl[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int c = c0 + k * BLOCK;
                pv[k] = p[c];
                vv[k] = v[c];
                rv[k] = r[c];
                if constexpr (!SWEEP) {
                    dgv[k] = sv.dg[c];
                    vl[k] = sv.vol[c];
                }
            }
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int c = c0 + k * BLOCK;
                const R pn = (pv[k] - (R)omega * vv[k]) * (R)beta + rv[k];
                p[c] = pn;
                if constexpr (!SWEEP) ph[c] = pn / (sv.vt * dgv[k] + sv.wv * (R)vl[k]);
            }
        });
        __syncthreads();
        if constexpr (SWEEP) {
            sweep<R, BLOCK>(sv, p, ph);
            __syncthreads();
        }
        // (2) v = A p^, <r_hat, v>
        red[0] = 0;
        for_cells<BLOCK, UM>(N, [&](auto kc, int c0) {
            constexpr int K = decltype(kc)::value;
            R rhv[K];
#pragma unroll
            for (int k = 0; k < K; k++) rhv[k] = rh[c0 + k * BLOCK];
            spmv_cells<R, BLOCK, K, !SWEEP>(sv, ph, p, v, c0,
                                            [&](int k, int, R vc) { red[0] += (double)rhv[k] * (double)vc; });
        });
        block_sum<BLOCK, 1>(red, lds);
        const double rv = red[0];
        if (rv < 1e-200 && rv > -1e-200) {
            // restart (bicgstab.hpp:224-235)
            rho = alpha = omega = 1;
            for (int c = threadIdx.x; c < N; c += BLOCK) {
                const R bc = rhs[c];
                x[c] = 0;
                v[c] = 0;
                p[c] = 0;
                r[c] = bc;
                rh[c] = refsem ? (R)(bc * (R)6.8 + (R)0.6) : bc;
            }
            __syncthreads();
            red[0] = 0;
            for (int c = threadIdx.x; c < N; c += BLOCK) red[0] += (double)rh[c] * (double)r[c];
            block_sum<BLOCK, 1>(red, lds);
            rho_next = red[0];
            continue;
        }
        alpha = rho / rv;
        // (3) s = r - alpha v (stored in r), ||s||^2; Jacobi: s^ = s / diag
        red[0] = 0;
        for_cells<BLOCK, US>(N, [&](auto kc, int c0) {
            constexpr int K = decltype(kc)::value;
            R rv[K], vv[K], dgv[K];
            double vl[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int c = c0 + k * BLOCK;
                rv[k] = r[c];
                vv[k] = v[c];
                if constexpr (!SWEEP) {
                    dgv[k] = sv.dg[c];
                    vl[k] = sv.vol[c];
                }
            }
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int c = c0 + k * BLOCK;
                const R s = rv[k] - (R)alpha * vv[k];
                r[c] = s;
                if constexpr (!SWEEP) sh[c] = s / (sv.vt * dgv[k] + sv.wv * (R)vl[k]);
                red[0] += (double)s * (double)s;
            }
        });
        block_sum<BLOCK, 1>(red, lds);
        if (!refsem && sqrt(red[0]) < tol * nrm0) {
            for (int c = threadIdx.x; c < N; c += BLOCK) x[c] += (R)alpha * ph[c];
            __syncthreads();
            converged = true;
            break;
        }
        if constexpr (SWEEP) {
            sweep<R, BLOCK>(sv, r, sh);
            __syncthreads();
        }
        // (4) t = A s^, <t, s>, <t, t>
        red[0] = 0;
        red[1] = 0;
        for_cells<BLOCK, UM>(N, [&](auto kc, int c0) {
            constexpr int K = decltype(kc)::value;
            R sv_[K];
#pragma unroll
            for (int k = 0; k < K; k++) sv_[k] = r[c0 + k * BLOCK];
            spmv_cells<R, BLOCK, K, !SWEEP>(sv, sh, r, t, c0, [&](int k, int, R tc) {
                red[0] += (double)tc * (double)sv_[k];
                red[1] += (double)tc * (double)tc;
            });
        });
        block_sum<BLOCK, 2>(red, lds);
        omega = (red[1] != 0) ? red[0] / red[1] : 0.0;
        // (5) x += alpha p^ + omega s^, r = s - omega t, ||r||^2, <r_hat, r>
        red[0] = 0;
        red[1] = 0;
        for_cells<BLOCK, US>(N, [&](auto kc, int c0) {
            constexpr int K = decltype(kc)::value;
            R xv[K], phv[K], shv[K], rv[K], tv[K], rhv[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int c = c0 + k * BLOCK;
                xv[k] = x[c];
                phv[k] = ph[c];
                shv[k] = sh[c];
                rv[k] = r[c];
                tv[k] = t[c];
                rhv[k] = rh[c];
            }
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int c = c0 + k * BLOCK;
                x[c] = xv[k] + ((R)alpha * phv[k] + (R)omega * shv[k]);
                const R rn = rv[k] - (R)omega * tv[k];
                r[c] = rn;
                red[0] += (double)rn * (double)rn;
                red[1] += (double)rhv[k] * (double)rn;
            }
        });
        block_sum<BLOCK, 2>(red, lds);
        nrm = sqrt(red[0]);
        rho_next = red[1];
        if (nrm / nrm0 < tol) {
            converged = true;
            break;
        }
        // stagnation: no new smallest residual for stall_window iterations
        if (nrm < best) {
            best = nrm;
            best_it = i;
        } else if (stall_window > 0 && i - best_it >= stall_window) {
            break;
        }
    }
    __syncthreads();
    return converged ? it : maxit + 1;
}

template <typename R>
__device__ __forceinline__ SysView<R> make_view(const Ops& o, int b, int d);

template <>
__device__ __forceinline__ SysView<double> make_view<double>(const Ops& o, int b, int d) {
    SysView<double> sv;
    sv.N = o.N;
    sv.W = o.W;
    sv.d = d;
    sv.vt = o.vt[b];
    sv.wv = o.omega[d];
    sv.vol = o.vol;
    sv.dg = o.dg + (size_t)d * o.N;
    sv.col = o.ell_col + (size_t)d * o.W * o.N;
    sv.val = o.ell_val + (size_t)d * o.W * o.N;
    sv.L = o.lvl_cnt ? o.lvl_cnt[d] : 0;
    sv.lp = o.lvl_ptr ? o.lvl_ptr + o.lvl_off[d] : nullptr;
    sv.perm = o.perm ? o.perm + (size_t)d * o.N : nullptr;
    sv.scol = o.swp_col ? o.swp_col + (size_t)d * o.W * o.N : nullptr;
    sv.sval = o.swp_val ? o.swp_val + (size_t)d * o.W * o.N : nullptr;
    sv.sdg = o.swp_dg ? o.swp_dg + (size_t)d * o.N : nullptr;
    return sv;
}

template <>
__device__ __forceinline__ SysView<float> make_view<float>(const Ops& o, int b, int d) {
    SysView<float> sv;
    sv.N = o.N;
    sv.W = o.W;
    sv.d = d;
    sv.vt = (float)o.vt[b];
    sv.wv = (float)o.omega[d];
    sv.vol = o.vol;
    sv.dg = o.dg32 + (size_t)d * o.N;
    sv.col = o.ell_col + (size_t)d * o.W * o.N;
    sv.val = o.ell_val32 + (size_t)d * o.W * o.N;
    sv.L = o.lvl_cnt ? o.lvl_cnt[d] : 0;
    sv.lp = o.lvl_ptr ? o.lvl_ptr + o.lvl_off[d] : nullptr;
    sv.perm = o.perm ? o.perm + (size_t)d * o.N : nullptr;
    sv.scol = o.swp_col ? o.swp_col + (size_t)d * o.W * o.N : nullptr;
    sv.sval = o.swp_val32 ? o.swp_val32 + (size_t)d * o.W * o.N : nullptr;
    sv.sdg = o.swp_dg32 ? o.swp_dg32 + (size_t)d * o.N : nullptr;
    return sv;
}

// Exact inner solve by one downwind sweep with the right-hand side formed on the fly.  Valid when
// the level schedule lagged no upwind edge: the operator is then triangular in level order and the
// sweep is its exact inverse (what BiCGSTAB with the sweep preconditioner returns after its first
// step).  One work-group per system; ee_curr is written once per cell, no Krylov vectors.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sweep_direct(Geo g, Quad q, Bnd bd, Bands bands, Ops o, Work w, int S,
                                                       const double* __restrict__ ee_prev,
                                                       const double* __restrict__ dens) {
    for (int s = blockIdx.x; s < S; s += gridDim.x) {
        const int b = s / o.D, d = s % o.D;
        const SysView<double> sv = make_view<double>(o, b, d);
        double* x = w.x + (size_t)s * o.N;
        for (int l = 0; l < sv.L; l++) {
            const int beg = sv.lp[l], end = sv.lp[l + 1];
            for (int i = beg + (int)threadIdx.x; i < end; i += BLOCK) {
                const int c = sv.perm[i];
                const double rhs = rhs_at(g, q, bd, bands, b, d, c, ee_prev, dens);
                double acc = 0;
#pragma unroll 3
                for (int k = 0; k < sv.W; k++) {
                    const int j = sv.scol[(size_t)k * sv.N + i];
                    if (j >= 0) acc += sv.sval[(size_t)k * sv.N + i] * x[j];
                }
                x[c] = sweep_div(rhs - sv.vt * acc, diag_of(sv, sv.sdg[i], c));
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) w.iters[s] += 1;
    }
}

// Direct sweep, tuned form.  Per level every thread loads its cell's level-ordered data (operator
// slots, w_d V_c, boundary flag) contiguously in i; only the density and the upwind values are
// gathered by cell index.  Interior cells skip the face loop of the right-hand side.  With
// XLDS the system's intensity vector lives in LDS during the sweep (gathers and stores never
// leave the CU) and is written to HBM once, coalesced, at the end; this needs N*8 bytes of LDS.
template <int BLOCK, bool XLDS>
__device__ __forceinline__ void sweep_fast_system(const Geo& g, const Quad& q, const Bnd& bd, const Bands& bands,
                                                  const Ops& o, const Work& w, int s, double* xs,
                                                  const double* __restrict__ ee_prev,
                                                  const double* __restrict__ dens) {
    const int N = o.N, W = o.W;
    {
        const int b = s / o.D, d = s % o.D;
        const double vt = o.vt[b];
        double* xg = w.x + (size_t)s * N;
        double* x = XLDS ? xs : xg;
        const int* __restrict__ perm = o.perm + (size_t)d * N;
        const int* __restrict__ scol = o.swp_col + (size_t)d * W * N;
        const double* __restrict__ sval = o.swp_val + (size_t)d * W * N;
        const double* __restrict__ sdg = o.swp_dg + (size_t)d * N;
        const double* __restrict__ svw = o.swp_vw + (size_t)d * N;
        const unsigned char* __restrict__ sbnd = o.swp_bnd + (size_t)d * N;
        const double* __restrict__ densb = dens + b;  // cell-major [N][B]: cell c at densb[c * B]
        const int* lp = o.lvl_ptr + o.lvl_off[d];
        const int L = o.lvl_cnt[d];
        for (int l = 0; l < L; l++) {
            const int beg = lp[l], end = lp[l + 1];
            for (int i = beg + (int)threadIdx.x; i < end; i += BLOCK) {
                const int c = perm[i];
                const double vw = svw[i];
                const double dgc = sdg[i];
                int j[3];
                double a[3];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    j[k] = (k < W) ? scol[(size_t)k * N + i] : -1;
                    a[k] = (k < W) ? sval[(size_t)k * N + i] : 0.0;
                }
                const double rhs =
                    sbnd[i] ? rhs_at(g, q, bd, bands, b, d, c, ee_prev, dens) : densb[(size_t)c * bands.B] * vw;
                double acc = 0;
#pragma unroll
                for (int k = 0; k < 3; k++)
                    if (j[k] >= 0) acc += a[k] * x[j[k]];
                x[c] = sweep_div(rhs - vt * acc, vt * dgc + vw);
            }
            __syncthreads();
        }
        if constexpr (XLDS) {
            for (int c = threadIdx.x; c < N; c += BLOCK) xg[c] = xs[c];
            __syncthreads();
        }
        if (threadIdx.x == 0) w.iters[s] += 1;
    }
}

template <int BLOCK, bool XLDS>
__global__ __launch_bounds__(BLOCK) void k_sweep_fast(Geo g, Quad q, Bnd bd, Bands bands, Ops o, Work w, int S,
                                                     const double* __restrict__ ee_prev,
                                                     const double* __restrict__ dens) {
    extern __shared__ double xs[];
    for (int s = blockIdx.x; s < S; s += gridDim.x)
        sweep_fast_system<BLOCK, XLDS>(g, q, bd, bands, o, w, s, xs, ee_prev, dens);
}

// Right-hand sides of every system in sweep (level) order, rhs_lo[s*Np + i] for cell perm_d[i]: the
// emission term for interior cells, rhs_at for cells with an incoming boundary face -- the values
// the level sweeps form on the fly, computed up front, fully parallel, so that the pipelined sweep
// streams them with its other operands instead of gathering by cell inside the level chain.
// One thread per level position (d, i), for all bands: the cell's densities are one contiguous
// record (dens is cell-major, [N][B]), so a wave's gather touches each line once for all bands, and
// for each band consecutive threads store consecutive positions of one system (coalesced).  The
// record is read in chunks of up to 12 bands, each chunk's band count a compile-time constant (a
// uniform switch), so that its loads are all issued before its stores (loads and stores interleaved
// make every store wait for the next load).  Cells with an incoming boundary face get the interior
// value too: k_rhs_levels_bnd, which runs after this pass, overwrites it.
template <int NB>
__device__ __forceinline__ void rhs_band_chunk(const double* __restrict__ dc, double vw, double* __restrict__ out,
                                               size_t stride) {
    double v[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) v[k] = dc[k];
#pragma unroll
    for (int k = 0; k < NB; k++) out[(size_t)k * stride] = v[k] * vw;
}

// The same from the source temperature t of the cell: C_b / solid_angle * (t - T_ref) is the
// emission density as its writers form it, so the products are the same.
template <int NB>
__device__ __forceinline__ void rhs_band_chunk_t(const double* __restrict__ C, double solid_angle, double tm, double vw,
                                                 double* __restrict__ out, size_t stride) {
#pragma unroll
    for (int k = 0; k < NB; k++) out[(size_t)k * stride] = C[k] / solid_angle * tm * vw;
}

// tsrc (or null): the source temperatures, whose 8 bytes per cell the pass gathers instead of the
// cell's B densities (the gather by cell is the pass's cost; the temperatures stay in L2).
__device__ __forceinline__ void rhs_levels_range(const Ops& o, int B, const double* __restrict__ dens,
                                                 const double* __restrict__ tsrc, double* __restrict__ rhs_lo,
                                                 int blk, int nblk) {
    const int N = o.N, D = o.D, Np = o.Np;
    const long long total = (long long)D * Np;
    const size_t stride = (size_t)D * Np;  // system (b, d) -> (b + 1, d)
    for (long long t = blk * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)nblk * blockDim.x) {
        const int d = (int)(t / Np), i = (int)(t % Np);
        if (i >= N) continue;                // padding of the row
        const size_t u = (size_t)d * N + i;  // the direction's level-order operands are unpadded
        const double vw = o.swp_vw[u];
        if (tsrc) {
            const double tm = tsrc[o.perm[u]] - o.T_ref;
            for (int b0 = 0; b0 < B; b0 += 12) {
                const double* c0 = o.band_C + b0;
                double* o0 = rhs_lo + t + (size_t)b0 * stride;
                switch (min(12, B - b0)) {
                    case 1: rhs_band_chunk_t<1>(c0, o.solid_angle, tm, vw, o0, stride); break;
                    case 2: rhs_band_chunk_t<2>(c0, o.solid_angle, tm, vw, o0, stride); break;
                    case 3: rhs_band_chunk_t<3>(c0, o.solid_angle, tm, vw, o0, stride); break;
                    case 4: rhs_band_chunk_t<4>(c0, o.solid_angle, tm, vw, o0, stride); break;
                    case 5: rhs_band_chunk_t<5>(c0, o.solid_angle, tm, vw, o0, stride); break;
                    case 6: rhs_band_chunk_t<6>(c0, o.solid_angle, tm, vw, o0, stride); break;
                    case 7: rhs_band_chunk_t<7>(c0, o.solid_angle, tm, vw, o0, stride); break;
                    case 8: rhs_band_chunk_t<8>(c0, o.solid_angle, tm, vw, o0, stride); break;
                    case 9: rhs_band_chunk_t<9>(c0, o.solid_angle, tm, vw, o0, stride); break;
                    case 10: rhs_band_chunk_t<10>(c0, o.solid_angle, tm, vw, o0, stride); break;
                    case 11: rhs_band_chunk_t<11>(c0, o.solid_angle, tm, vw, o0, stride); break;
                    default: rhs_band_chunk_t<12>(c0, o.solid_angle, tm, vw, o0, stride); break;
                }
            }
            continue;
        }
        const double* __restrict__ dc = dens + (size_t)o.perm[u] * B;
        for (int b0 = 0; b0 < B; b0 += 12) {
            const double* c0 = dc + b0;
            double* o0 = rhs_lo + t + (size_t)b0 * stride;
            switch (min(12, B - b0)) {
                case 1: rhs_band_chunk<1>(c0, vw, o0, stride); break;
                case 2: rhs_band_chunk<2>(c0, vw, o0, stride); break;
                case 3: rhs_band_chunk<3>(c0, vw, o0, stride); break;
                case 4: rhs_band_chunk<4>(c0, vw, o0, stride); break;
                case 5: rhs_band_chunk<5>(c0, vw, o0, stride); break;
                case 6: rhs_band_chunk<6>(c0, vw, o0, stride); break;
                case 7: rhs_band_chunk<7>(c0, vw, o0, stride); break;
                case 8: rhs_band_chunk<8>(c0, vw, o0, stride); break;
                case 9: rhs_band_chunk<9>(c0, vw, o0, stride); break;
                case 10: rhs_band_chunk<10>(c0, vw, o0, stride); break;
                case 11: rhs_band_chunk<11>(c0, vw, o0, stride); break;
                default: rhs_band_chunk<12>(c0, vw, o0, stride); break;
            }
        }
    }
}

__global__ void k_rhs_levels(Ops o, int B, const double* __restrict__ dens, const double* __restrict__ tsrc,
                             double* __restrict__ rhs_lo) {
    rhs_levels_range(o, B, dens, tsrc, rhs_lo, blockIdx.x, gridDim.x);
}

// Source temperatures in level order, tlo[d * Np + i] = tsrc[cell at position i of direction d],
// for the sweeps that form the interior right-hand sides themselves: 8 bytes per position shared by
// all bands, against the B right-hand sides k_rhs_levels stores.
__device__ __forceinline__ void tsrc_levels_range(const Ops& o, const double* __restrict__ tsrc,
                                                  double* __restrict__ tlo, int blk, int nblk) {
    const int N = o.N, Np = o.Np;
    const long long total = (long long)o.D * Np;
    for (long long t = blk * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)nblk * blockDim.x) {
        const int d = (int)(t / Np), i = (int)(t % Np);
        if (i < N) tlo[t] = tsrc[o.perm[(size_t)d * N + i]];
    }
}

__global__ void k_tsrc_levels(Ops o, const double* __restrict__ tsrc, double* __restrict__ tlo) {
    tsrc_levels_range(o, tsrc, tlo, blockIdx.x, gridDim.x);
}

// The diffuse re-emission (work-groups [0, n_diff)) and the level-order source temperatures.
__global__ void k_diffuse_tsrc_levels(Geo g, Quad q, Bnd bd, Ops o, int B, const double* __restrict__ ee_prev,
                                      double* __restrict__ ediff, const double* __restrict__ tsrc,
                                      double* __restrict__ tlo, int n_diff) {
    if ((int)blockIdx.x < n_diff)
        diffuse_range(g, q, bd, B, ee_prev, ediff, blockIdx.x, n_diff);
    else
        tsrc_levels_range(o, tsrc, tlo, blockIdx.x - n_diff, gridDim.x - n_diff);
}

// The diffuse re-emission (work-groups [0, n_diff)) and the interior level-order right-hand sides
// (the rest) in one launch: the two are independent, so the small diffuse pass runs beside the
// larger one.
__global__ void k_diffuse_rhs_levels(Geo g, Quad q, Bnd bd, Ops o, int B, const double* __restrict__ ee_prev,
                                     double* __restrict__ ediff, const double* __restrict__ dens,
                                     const double* __restrict__ tsrc, double* __restrict__ rhs_lo, int n_diff) {
    if ((int)blockIdx.x < n_diff)
        diffuse_range(g, q, bd, B, ee_prev, ediff, blockIdx.x, n_diff);
    else
        rhs_levels_range(o, B, dens, tsrc, rhs_lo, blockIdx.x - n_diff, gridDim.x - n_diff);
}

// The boundary cells of k_rhs_levels from a compact list of level-order indices t = d*N + i, so
// their face loops run in work-groups of their own instead of stalling the interior waves; one
// thread per (list entry, band).  A work-group takes blockDim.x consecutive entries of one band (its
// stores coalesced); the B work-groups of the same entries follow one another on one XCD (virtual
// work-group v runs on XCD v % 8 when the grid is a multiple of 8), so the cells' geometry,
// densities and incoming intensities they all read come from that XCD's L2 after the first band.
__host__ __device__ inline long long rhs_bnd_units(int n_list, int B, int E) {
    const long long chunks = (n_list + E - 1) / E;
    return 8 * (long long)B * ((chunks + 7) / 8);
}

template <class BndT>
__device__ __forceinline__ void rhs_bnd_range(const Geo& g, const Quad& q, const BndT& bd, const Bands& bands,
                                              const Ops& o, int S, const int* __restrict__ list, int n_list,
                                              const double* __restrict__ ee_prev, const double* __restrict__ dens,
                                              double* __restrict__ rhs_lo, int blk, int nblk) {
    const int N = o.N, D = o.D, B = S / D, E = blockDim.x;
    const long long V = rhs_bnd_units(n_list, B, E);
    for (long long v = blk; v < V; v += nblk) {
        const long long r = v / 8;
        const int b = (int)(r % B);
        const long long k = ((r / B) * 8 + v % 8) * E + threadIdx.x;
        if (k >= n_list) continue;
        const int t = list[k];
        const int d = t / N, i = t % N;
        rhs_lo[((size_t)b * D + d) * o.Np + i] = rhs_at(g, q, bd, bands, b, d, o.perm[t], ee_prev, dens);
    }
}

__global__ void k_rhs_levels_bnd(Geo g, Quad q, Bnd bd, Bands bands, Ops o, int S, const int* __restrict__ list,
                                 int n_list, const double* __restrict__ ee_prev, const double* __restrict__ dens,
                                 double* __restrict__ rhs_lo) {
    rhs_bnd_range(g, q, bd, bands, o, S, list, n_list, ee_prev, dens, rhs_lo, blockIdx.x, gridDim.x);
}

// Global-memory (HBM) views of pointers.  A pointer the kernel reads from memory (the uneven
// batch's per-problem descriptors) is generic to the compiler: loads through it become flat loads,
// which it counts as LDS traffic too, so every LDS wait of the pipelined sweeps would also wait for
// the operand prefetches.
template <class T>
using gmem = __attribute__((address_space(1))) T;
template <class T>
__device__ __forceinline__ gmem<T>* to_global(T* p) {
    return (gmem<T>*)p;
}

// Descriptors read through the constant address space: a pointer loaded from there is taken to be a
// global one (only the host fills such memory), so the loads and stores through it are global
// ones, not flat ones (which also count as LDS traffic and serialise the waits).
template <class T>
using cmem = __attribute__((address_space(4))) const T;
template <class T>
__device__ __forceinline__ T cload(cmem<T>* p) {
    T v;
    __builtin_memcpy(&v, (const T*)p, sizeof(T));
    return v;
}

// Boundary data of a descriptor in constant memory, its per-condition tables read in place (a copy
// of the tables, indexed at run time, would live in scratch memory).
struct BndView {
    int nbc;
    cmem<int>* bci;
    cmem<int>* bct;
    cmem<double>* bcT;
    int n_bf;
    const int* bf_cell;
    const int* bf_face;
    const int* bfid;
    const int* mirror;
    const double* ediff;
    int has_diffuse;
};
__device__ __forceinline__ BndView bnd_view(cmem<Bnd>* p) {
    return BndView{p->nbc, p->bci, p->bct, p->bcT, p->n_bf, p->bf_cell, p->bf_face, p->bfid, p->mirror, p->ediff,
                   p->has_diffuse};
}

// Raw buffer loads and stores: a buffer resource (four scalar registers) holds an array's base and
// size, a scalar byte offset selects a slot (band, operator column), and one vector byte offset per
// lane is shared by every array indexed alike -- in place of a 64-bit vector address per array.
typedef unsigned int bte_u32x2 __attribute__((ext_vector_type(2)));

// Raw buffer resource over [p, p + bytes) (stride 0; word 3 as for every gfx9 target).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void* p, size_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)(unsigned)bytes, 0x00020000);
}

// AUX: the cache-policy bits of the load (2: non-temporal on gfx950, for data read once).
template <typename T, int AUX = 0>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, AUX));
}

template <typename T>
__device__ __forceinline__ void buf_st(T v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(bte_u32x2, v), r, voff, soff, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}

// Operands of one cell of the pipelined sweeps from the packed records (k_pack_sweep).
struct PkItem {
    int i, c;     // level-order position (-1: no cell for this thread) and cell
    bool bnd;     // the cell has an incoming boundary face (loaded by the TS sweeps only)
    int p[3];     // 8 x level-order positions of the upwind neighbours (absent: INT_MAX)
    double a[3];  // their coefficients (absent: 0)
    double dgc, vw, rhs;
    double ts;    // source temperature (sweeps that form interior right-hand sides themselves)
};

struct PkSrc {
    __amdgpu_buffer_rsrc_t pki, pkd, scol, svw, rhs, x, tlo, sbnd;
    unsigned n4;   // byte distance between the upwind-cell slots of swp_col: N * 4
    unsigned xs8;  // byte distance between the intensities of neighbouring cells in x (o.ee.sc * 8)
};

template <int W>
__device__ __forceinline__ PkSrc pk_src(const Ops& o, const Work& w, int s, int d) {
    const size_t N = (size_t)o.N;
    PkSrc r;
    r.pki = buffer_of(o.swp_pki + d * N * 4, N * 16);
    r.pkd = buffer_of(o.swp_pkd + d * N * 4, N * 32);
    r.scol = buffer_of(o.swp_col + d * W * N, W * N * 4);
    r.svw = buffer_of(o.swp_vw + d * N, N * 8);
    r.rhs = buffer_of(w.b + (size_t)s * o.Np, N * 8);
    r.tlo = buffer_of(w.tlo ? w.tlo + d * (size_t)o.Np : w.b, N * 8);
    r.sbnd = buffer_of(o.swp_bnd + d * N, N);
    const long long sc = o.ee.sc;
    r.x = buffer_of(w.x + o.ee.at(s / o.D, d, 0), (size_t)((N - 1) * sc + 1) * 8);
    r.n4 = (unsigned)(N * 4);
    r.xs8 = (unsigned)(sc * 8);
    return r;
}

typedef unsigned int bte_u32x4 __attribute__((ext_vector_type(4)));
typedef double bte_f64x2 __attribute__((ext_vector_type(2)));

// Three 16-byte loads (four with W = 3) and the right-hand side, all at the lane's position; an
// absent item (i < 0) loads position 0 and is ignored.  TS: also the source temperature (shared
// by the bands of the direction, so a plain cached load) and whether the cell has an incoming
// boundary face (swp_bnd).
template <int W, bool RHS = true, bool TS = false>
__device__ __forceinline__ void pk_item_load(int i, const PkSrc& src, PkItem& it) {
    it.i = i;
    const unsigned j = (unsigned)(i < 0 ? 0 : i);
    const bte_u32x4 q = __builtin_amdgcn_raw_buffer_load_b128(src.pki, j * 16u, 0, 0);
    const bte_f64x2 d0 = __builtin_bit_cast(bte_f64x2, __builtin_amdgcn_raw_buffer_load_b128(src.pkd, j * 32u, 0, 0));
    const bte_f64x2 d1 =
        __builtin_bit_cast(bte_f64x2, __builtin_amdgcn_raw_buffer_load_b128(src.pkd, j * 32u + 16u, 0, 0));
    it.c = (int)q[0];
    it.p[0] = (int)q[1];
    it.p[1] = (int)q[2];
    it.p[2] = (int)q[3];
    if constexpr (W == 1) {
        it.a[0] = d0[0];
        it.a[1] = it.a[2] = 0.0;
        it.dgc = d0[1];
        it.vw = d1[0];
    } else if constexpr (W == 2) {
        it.a[0] = d0[0];
        it.a[1] = d0[1];
        it.a[2] = 0.0;
        it.dgc = d1[0];
        it.vw = d1[1];
    } else {
        it.a[0] = d0[0];
        it.a[1] = d0[1];
        it.a[2] = d1[0];
        it.dgc = d1[1];
        it.vw = buf_ld<double>(src.svw, j * 8u, 0);
    }
    if constexpr (RHS) it.rhs = buf_ld<double, 2>(src.rhs, j * 8u, 0);
    if constexpr (TS) {
        it.ts = buf_ld<double>(src.tlo, j * 8u, 0);
        it.bnd = __builtin_amdgcn_raw_buffer_load_b8(src.sbnd, j, 0, 0) != 0;
    }
}

// Upwind sum of a pipelined-sweep cell, sum_k a_k x(p_k) in slot order over the W slots (the
// order of the level sweep): positions p >= end - R from the LDS ring, older ones from HBM.  An
// absent slot reads the ring's last slot (finite: cleared before the sweep, then sweep values)
// times 0, which leaves the sum unchanged.  HB: the level has reads older than the ring, resolved
// after one wave-uniform test; the upwind cell of such a read is loaded only then, and its value
// is consumed (a register move) inside its own branch: a load still pending where the branches
// merge makes the compiler wait there for every outstanding load, the operand prefetches of later
// levels included (a select between the two sources is worse: the compiler merges them into one
// flat load and waits for all loads after every cell).
// Positions are byte offsets (8 p): the ring slot is p8 & (8 R - 8), and lim8 = 8 (end - R).
template <int W, bool HB>
__device__ __forceinline__ double pk_upwind_sum(const PkItem& it, int lim8, const double* ring, int mask8,
                                                const gmem<double>* xg, const PkSrc& src) {
    double v[W];
#pragma unroll
    for (int k = 0; k < W; k++) v[k] = *(const double*)((const char*)ring + (it.p[k] & mask8));
    if constexpr (HB) {
        bool old = false;
#pragma unroll
        for (int k = 0; k < W; k++) old |= it.p[k] < lim8;
        if (__ballot(old)) {
#pragma unroll
            for (int k = 0; k < W; k++)
                if (it.p[k] < lim8) {
                    const int c = buf_ld<int>(src.scol, (unsigned)it.i * 4u, k * src.n4);
                    const double h = xg[(size_t)c * (src.xs8 >> 3)];
                    asm volatile("v_mov_b64 %0, %1" : "=v"(v[k]) : "v"(h));
                }
        }
    }
    double acc = 0;
#pragma unroll
    for (int k = 0; k < W; k++) acc += it.a[k] * v[k];
    return acc;
}

// LDS-only work-group barrier: the ring is the only memory threads exchange inside the sweep, so
// outstanding HBM loads (the prefetches) stay in flight across it.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// LDS of the pipelined sweep: the ring, then the level pointers of the system's direction.
__host__ __device__ inline size_t ring_lds_bytes(int ring, int max_levels) {
    return (size_t)ring * sizeof(double) + (size_t)(max_levels + 1) * sizeof(int);
}

// One system of the pipelined sweep; needs rhs_lo (k_rhs_levels) in w.b.  ring holds the LDS ring
// of R = ring_mask + 1 values followed by room for the direction's level pointers.  Per thread every
// operand of a level -- including its right-hand side -- is a coalesced load in level order, issued
// PD levels ahead.  An upwind value at position p >= end(l) - R is read from the ring (its slot is
// not yet reused).  An older one is read from HBM: the host flags (o.lvl_sync) the levels after
// which the work-group synchronises with its HBM writes complete, chosen so that one such
// synchronisation lies between the write of every value read this way and its read.  All other
// levels synchronise on the LDS only, leaving the prefetches in flight.  FB = false: no direction
// of the launch has such reads (no HBM read-back path, no full synchronisation); with FB = true a
// direction without them (its o.dir_fb bit clear) still runs the FB = false code.  o.ring_sel picks
// the flags computed for the launch's ring (the regular or the wide one).  TS: w.b holds the
// right-hand sides of the cells with an incoming boundary face only; an interior cell's is formed
// here from its source temperature (w.tlo), (C_b / solid_angle * (T - T_ref)) * w_d V_c -- the
// emission density times the volume weight, rounded as k_rhs_levels forms it.
template <int BLOCK, int W, bool FB, int PD = 3, bool TS = false>
__device__ __forceinline__ void sweep_ring_system(const Ops& o, const Work& w, int s, double* ring, int ring_mask) {
    const int tid = threadIdx.x;
    const int R = ring_mask + 1;
    int* lps = (int*)(ring + R);
    {
        const int b = s / o.D, d = s % o.D;
        const double vt = o.vt[b];
        const double cb = TS ? o.band_C[b] / o.solid_angle : 0.0, t_ref = o.T_ref;
        gmem<double>* __restrict__ xg = to_global(w.x + o.ee.at(b, d, 0));
        const PkSrc src = pk_src<W>(o, w, s, d);
        const int* __restrict__ lp = o.lvl_ptr + o.lvl_off[d];
        const int L = o.lvl_cnt[d];
        // level pointers with the level's flags (o.lvl_sync) in bits 31 (synchronise after it) and
        // 30 (it reads values older than the ring)
        if constexpr (FB) {
            const unsigned char* __restrict__ fl = o.lvl_sync + o.lvl_off[d];
            const int sh = 2 * o.ring_sel;
            for (int t = tid; t <= L; t += BLOCK) lps[t] = lp[t] | (t < L ? (int)((fl[t] >> sh) & 3) << 30 : 0);
        } else {
            for (int t = tid; t <= L; t += BLOCK) lps[t] = lp[t];
        }
        if (tid == 0) ring[ring_mask] = 0.0;  // read, times 0, for absent upwind slots
        __syncthreads();
        // Level pointers lps[wb .. wb + 63] are held one per lane and read with readlane (a register
        // read, no LDS round trip on the level's dependency chain), each flag as one 64-bit mask;
        // the window moves before a level would need a pointer beyond it (level l reads l .. l+PD+1).
        const int lane = tid & 63;
        int wb = 0, win = 0;
        unsigned long long syncs = 0, olds = 0;
        // the flags hold for rings of at least o.flag_ring() values; a smaller ring treats every level
        // as flagged
        const bool every = R < o.flag_ring();
        auto load_window = [&](int from) {
            wb = from;
            const int v = lps[min(from + lane, L)];
            win = v & 0x3fffffff;
            if constexpr (FB) {
                syncs = every ? ~0ull : __ballot(v & 0x80000000);
                olds = every ? ~0ull : __ballot(v & 0x40000000);
            }
        };
        load_window(0);
        auto lp_at = [&](int l) { return __builtin_amdgcn_readlane(win, l - wb); };
        auto first = [&](int l) {  // this thread's first-pass position in level l, or -1
            if (l >= L) return -1;
            const int i = lp_at(l) + tid;
            return i < lp_at(l + 1) ? i : -1;
        };
        // HB: the level may read values older than the ring (HBM read-back path compiled in)
        auto solve_cell = [&](const PkItem& it, int end, auto hb) {
            const double acc =
                pk_upwind_sum<W, decltype(hb)::value>(it, (end - R) * 8, ring, ring_mask * 8, xg, src);
            double rhs = it.rhs;
            if constexpr (TS) rhs = it.bnd ? it.rhs : cb * (it.ts - t_ref) * it.vw;
            const double xv = sweep_div(rhs - vt * acc, vt * it.dgc + it.vw);
            ring[it.i & ring_mask] = xv;
            buf_st<double>(xv, src.x, (unsigned)it.c * src.xs8, 0);
        };
        auto solve_level = [&](const PkItem& cur, int beg, int end, auto hb) {
            if (cur.i >= 0) solve_cell(cur, end, hb);
            for (int i = beg + BLOCK + tid; i < end; i += BLOCK) {
                PkItem E;
                pk_item_load<W, true, TS>(i, src, E);
                solve_cell(E, end, hb);
            }
        };
        // level l computes from the item loaded PD levels earlier and loads the item of level l+PD
        // into the register set it just freed; the PD + 1 sets rotate by unrolling, never by copies
        // (a copy of a register with a load in flight would wait for that load)
        auto level = [&](int l, PkItem& cur, PkItem& ahead) {
            if (l + PD + 1 - wb > 63) load_window(l);
            const int beg = lp_at(l), end = lp_at(l + 1);
            pk_item_load<W, true, TS>(first(l + PD), src, ahead);
            if (FB && ((olds >> (l - wb)) & 1ull))
                solve_level(cur, beg, end, BoolC<true>{});
            else
                solve_level(cur, beg, end, BoolC<false>{});
            if (FB && ((syncs >> (l - wb)) & 1ull))
                __syncthreads();  // HBM writes of the levels so far complete and visible
            else
                lds_barrier();
        };
        PkItem X[PD + 1];
#pragma unroll
        for (int k = 0; k < PD; k++) pk_item_load<W, true, TS>(first(k), src, X[k]);
        X[PD].i = -1;
        // Steady state: every copy of the unrolled body runs, so on every path into a level the
        // compiler sees its item loaded PD levels earlier and waits for those loads only (a copy
        // that may be skipped inside the loop would make it wait for the last level's loads too).
        int l = 0;
        for (; l + PD + 1 <= L; l += PD + 1) {
#pragma unroll
            for (int k = 0; k <= PD; k++) level(l + k, X[k], X[(k + PD) % (PD + 1)]);
        }
#pragma unroll
        for (int k = 0; k < PD; k++)
            if (l + k < L) level(l + k, X[k], X[(k + PD) % (PD + 1)]);
        __syncthreads();  // the ring and level pointers are reused by the next system of this work-group
        if (tid == 0) w.iters[s] += 1;
    }
}

// FB = true: the HBM read-back path only for a direction that has such reads
template <int BLOCK, int W, bool FB, int PD = 3, bool TS = false>
__device__ __forceinline__ void sweep_ring_dispatch(const Ops& o, const Work& w, int s, double* ring, int ring_mask) {
    if (FB && (((o.dir_fb[s % o.D] >> o.ring_sel) & 1) || ring_mask + 1 < o.flag_ring()))
        sweep_ring_system<BLOCK, W, true, PD, TS>(o, w, s, ring, ring_mask);
    else
        sweep_ring_system<BLOCK, W, false, PD, TS>(o, w, s, ring, ring_mask);
}

// Accelerator complex dies (XCDs) of the MI355X; work-group blk of a launch runs on XCD blk % 8, each
// XCD with its own L2.
constexpr int kXcds = 8;

// System swept by work-group blk of a launch of S = B*D work-groups, one per system: XCD x sweeps
// the directions d = x, x + 8, ... and each direction's B bands in consecutive work-groups of that
// XCD, so the bands of a direction run at the same time, level by level, and all but the first
// read the direction's operator records from the XCD's L2.  The identity when D is not a multiple
// of 8.  A bijection of [0, S).
__device__ __forceinline__ int system_of_block(int blk, int B, int D) {
    if (D % kXcds != 0) return blk;
    const int x = blk % kXcds, r = blk / kXcds;
    return (r % B) * D + (r / B) * kXcds + x;
}

// W = incoming slots per cell (F - 1: 1 in 1D, 2 in 2D, 3 in 3D); TS: interior right-hand sides
// from the source temperatures (sweep_ring_system)
template <int BLOCK, int W, bool FB, bool TS = false, int PD = 3>
__global__ __launch_bounds__(BLOCK) void k_sweep_ring(Ops o, Work w, int S, int ring_mask) {
    extern __shared__ double ring[];
    for (int blk = blockIdx.x; blk < S; blk += gridDim.x)
        sweep_ring_dispatch<BLOCK, W, FB, PD, TS>(o, w, system_of_block(blk, S / o.D, o.D), ring, ring_mask);
}

// Band-batched pipelined sweep (variants 6 and 7): one work-group sweeps G bands of one direction.
// The bands share the level schedule and the stored operator, so each cell's operator item (cell,
// upwind positions and cells, coefficients, diagonal face sum, volume weight) is loaded once for
// all G bands, together with the G right-hand sides; each band keeps its own LDS ring of R values
// (the rings and then the level pointers fill the dynamic LDS).  Per band the arithmetic, and its
// order, are those of sweep_ring_system.  A band beyond the last one (a short last group) is
// skipped.
template <int BLOCK, int W, int G, bool FB>
__device__ __forceinline__ void sweep_ring_bands(const Ops& o, const Work& w, int B, int d, int b0, double* ring,
                                                 int ring_mask) {
    const int N = o.N;
    const int tid = threadIdx.x;
    const int R = ring_mask + 1;
    const int K = o.ring_sync;
    int* lps = (int*)(ring + G * R);
    const int nb = min(G, B - b0);  // bands present in this group
    double vt[G];
    gmem<double>* xg[G];
    __amdgpu_buffer_rsrc_t rhsr[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int s = (b0 + (g < nb ? g : 0)) * o.D + d;
        vt[g] = o.vt[b0 + (g < nb ? g : 0)];
        xg[g] = to_global(w.x + (size_t)s * N);
        rhsr[g] = buffer_of(w.b + (size_t)s * o.Np, (size_t)N * 8);
    }
    const PkSrc src = pk_src<W>(o, w, b0 * o.D + d, d);
    const int* __restrict__ lp = o.lvl_ptr + o.lvl_off[d];
    const int L = o.lvl_cnt[d];
    for (int t = tid; t <= L; t += BLOCK) lps[t] = lp[t];
    if (tid < G) ring[tid * R + ring_mask] = 0.0;  // read, times 0, for absent upwind slots
    __syncthreads();
    auto first = [&](int l) {  // this thread's first-pass position in level l, or -1
        if (l >= L) return -1;
        const int i = lps[l] + tid;
        return i < lps[l + 1] ? i : -1;
    };
    auto load = [&](int i, PkItem& it, double (&rhs)[G]) {
        pk_item_load<W, false>(i, src, it);
        const unsigned j8 = (unsigned)(i < 0 ? 0 : i) * 8u;
#pragma unroll
        for (int g = 0; g < G; g++) rhs[g] = buf_ld<double>(rhsr[g], j8, 0);
    };
    auto solve_cell = [&](const PkItem& it, const double (&rhs)[G], int end) {
#pragma unroll
        for (int g = 0; g < G; g++)
            if (g < nb) {
                double* rg = ring + g * R;
                const double acc = pk_upwind_sum<W, FB>(it, (end - R) * 8, rg, ring_mask * 8, xg[g], src);
                const double xv = sweep_div(rhs[g] - vt[g] * acc, vt[g] * it.dgc + it.vw);
                rg[it.i & ring_mask] = xv;
                xg[g][it.c] = xv;
            }
    };
    // as in sweep_ring_system: three register sets rotate by unrolling, never by copies
    auto level = [&](int l, PkItem& cur, double (&rcur)[G], PkItem& ahead, double (&rahead)[G]) {
        const int beg = lps[l], end = lps[l + 1];
        load(first(l + 2), ahead, rahead);
        if (cur.i >= 0) solve_cell(cur, rcur, end);
        for (int i = beg + BLOCK + tid; i < end; i += BLOCK) {
            PkItem E;
            double rE[G];
            load(i, E, rE);
            solve_cell(E, rE, end);
        }
        if (FB && (l + 1) % K == 0)
            __syncthreads();  // HBM writes of the levels so far complete and visible
        else
            lds_barrier();
    };
    PkItem X0, X1, X2;
    double r0[G], r1[G], r2[G];
    load(first(0), X0, r0);
    load(first(1), X1, r1);
    X2.i = -1;
    for (int l = 0; l < L; l += 3) {
        level(l, X0, r0, X2, r2);
        if (l + 1 < L) level(l + 1, X1, r1, X0, r0);
        if (l + 2 < L) level(l + 2, X2, r2, X1, r1);
    }
    __syncthreads();  // the rings and level pointers are reused by the next group of this work-group
    if (tid < nb) w.iters[(b0 + tid) * o.D + d] += 1;
}

template <int BLOCK, int W, int G, bool FB>
__global__ __launch_bounds__(BLOCK) void k_sweep_ring_bands(Ops o, Work w, int B, int ring_mask) {
    extern __shared__ double ring[];
    const int ngroups = o.D * ((B + G - 1) / G);
    for (int gid = blockIdx.x; gid < ngroups; gid += gridDim.x)
        sweep_ring_bands<BLOCK, W, G, FB>(o, w, B, gid % o.D, (gid / o.D) * G, ring, ring_mask);
}

// Single-wave form of the pipelined sweep (variant 5): one 64-lane wave per system, IPL cells of a
// level per lane, all loaded two levels ahead.  A wave's LDS instructions execute in order, so the
// ring needs no work-group barrier between levels: waiting for the wave's own LDS writes is enough.
// Values older than the ring come from HBM under the same rule as sweep_ring_system (after the
// levels the host flags the wave waits for its HBM writes).  Levels wider than 64 * IPL finish their
// extra cells unpipelined.  Same arithmetic, in the same order, as the other sweeps.
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

template <int W, int IPL, bool FB>
__device__ __forceinline__ void sweep_wave_system(const Ops& o, const Work& w, int s, double* ring, int ring_mask) {
    const int N = o.N;
    const int lane = threadIdx.x;
    const int R = ring_mask + 1;
    int* lps = (int*)(ring + R);
    const int b = s / o.D, d = s % o.D;
    const double vt = o.vt[b];
    gmem<double>* __restrict__ xg = to_global(w.x + (size_t)s * N);
    const PkSrc src = pk_src<W>(o, w, s, d);
    const int* __restrict__ lp = o.lvl_ptr + o.lvl_off[d];
    const int L = o.lvl_cnt[d];
    // level pointers and synchronisation flags as in sweep_ring_system
    if constexpr (FB) {
        const unsigned char* __restrict__ fl = o.lvl_sync + o.lvl_off[d];
        for (int t = lane; t <= L; t += 64) lps[t] = lp[t] | (t < L ? (int)(fl[t] & 3) << 30 : 0);
    } else {
        for (int t = lane; t <= L; t += 64) lps[t] = lp[t];
    }
    if (lane == 0) ring[ring_mask] = 0.0;  // read, times 0, for absent upwind slots
    __syncthreads();
    int wb = 0, win = 0;
    unsigned long long syncs = 0;
    auto load_window = [&](int from) {
        wb = from;
        const int v = lps[min(from + lane, L)];
        win = v & 0x3fffffff;
        if constexpr (FB) syncs = __ballot(v & 0x80000000);
    };
    load_window(0);
    const bool every = R < o.ring_r;
    auto lp_at = [&](int l) { return __builtin_amdgcn_readlane(win, l - wb); };
    auto first = [&](int l, int k) {
        if (l >= L) return -1;
        const int i = lp_at(l) + lane + 64 * k;
        return i < lp_at(l + 1) ? i : -1;
    };
    auto solve_cell = [&](const PkItem& it, int end) {
        const double acc = pk_upwind_sum<W, FB>(it, (end - R) * 8, ring, ring_mask * 8, xg, src);
        const double xv = sweep_div(it.rhs - vt * acc, vt * it.dgc + it.vw);
        ring[it.i & ring_mask] = xv;
        xg[it.c] = xv;
    };
    auto level = [&](int l, PkItem (&cur)[IPL], PkItem (&ahead)[IPL]) {
        if (l + 3 - wb > 63) load_window(l);
        const int beg = lp_at(l), end = lp_at(l + 1);
#pragma unroll
        for (int k = 0; k < IPL; k++)
            pk_item_load<W>(first(l + 2, k), src, ahead[k]);
#pragma unroll
        for (int k = 0; k < IPL; k++)
            if (cur[k].i >= 0) solve_cell(cur[k], end);
        for (int i = beg + 64 * IPL + lane; i < end; i += 64) {
            PkItem E;
            pk_item_load<W>(i, src, E);
            solve_cell(E, end);
        }
        if (FB && (every || ((syncs >> (l - wb)) & 1ull)))
            __syncthreads();  // HBM writes so far complete
        else
            wave_lds_sync();
    };
    PkItem X0[IPL], X1[IPL], X2[IPL];
#pragma unroll
    for (int k = 0; k < IPL; k++) {
        pk_item_load<W>(first(0, k), src, X0[k]);
        pk_item_load<W>(first(1, k), src, X1[k]);
        X2[k].i = -1;
    }
    // unconditional steady-state body, tail peeled (as in sweep_ring_system)
    int l = 0;
    for (; l + 3 <= L; l += 3) {
        level(l, X0, X2);
        level(l + 1, X1, X0);
        level(l + 2, X2, X1);
    }
    if (l < L) level(l, X0, X2);
    if (l + 1 < L) level(l + 1, X1, X0);
    __syncthreads();
    if (lane == 0) w.iters[s] += 1;
}

template <int W, bool FB>
__global__ __launch_bounds__(64) void k_sweep_wave(Ops o, Work w, int S, int ring_mask) {
    extern __shared__ double ring[];
    for (int s = blockIdx.x; s < S; s += gridDim.x) {
        if (FB && ((o.dir_fb[s % o.D] & 1) || ring_mask + 1 < o.ring_r))
            sweep_wave_system<W, 2, true>(o, w, s, ring, ring_mask);
        else
            sweep_wave_system<W, 2, false>(o, w, s, ring, ring_mask);
    }
}

// ----------------------------------------------------------------------------------------------
// Uneven batch of problems: one launch sweeps the systems of several meshes.  desc[k] holds problem
// k's data; items[t] = (problem, system) in the order the work-groups take them (longest level
// chains first).
// ----------------------------------------------------------------------------------------------
struct SweepDesc {
    Geo g;
    Quad q;
    Bnd bd;
    Bands bands;
    Ops o;
    Work w;
    const double* ee_prev;
    const double* dens;
};

template <int BLOCK, bool RING>
__global__ __launch_bounds__(BLOCK) void k_sweep_batch(const SweepDesc* __restrict__ desc, const int2* __restrict__ items,
                                                      int n_items, int ring_mask) {
    extern __shared__ double lds_x[];
    for (int t = blockIdx.x; t < n_items; t += gridDim.x) {
        const int2 it = items[t];
        if (it.x < 0) continue;  // padding of an XCD's sequence
        const SweepDesc& D = desc[it.x];
        if constexpr (RING)
            switch (D.o.W) {
                case 1: sweep_ring_dispatch<BLOCK, 1, true>(D.o, D.w, it.y, lds_x, ring_mask); break;
                case 2: sweep_ring_dispatch<BLOCK, 2, true>(D.o, D.w, it.y, lds_x, ring_mask); break;
                default: sweep_ring_dispatch<BLOCK, 3, true>(D.o, D.w, it.y, lds_x, ring_mask); break;
            }
        else
            sweep_fast_system<BLOCK, false>(D.g, D.q, D.bd, D.bands, D.o, D.w, it.y, lds_x, D.ee_prev, D.dens);
    }
}

// out[k] = *ptr[k]: the margins of the problems of a batch, for one copy to the host.
__global__ void k_gather_scalars(int n, const double* const* __restrict__ ptr, double* __restrict__ out) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) out[k] = *ptr[k];
}

// FP64 path: BiCGSTAB in FP64 on the right-hand sides in w.b (k_rhs) writes ee_curr directly.
template <int BLOCK, bool SWEEP>
__global__ __launch_bounds__(BLOCK) void k_solve64(Ops o, Work w, SolveCfg cfg) {
    __shared__ double lds[2 * (BLOCK / 64)];
    for (int s = blockIdx.x; s < cfg.S; s += gridDim.x) {
        const int b = s / o.D, d = s % o.D;
        const size_t base = (size_t)s * o.N;
        const double* rhs = w.b + base;
        const SysView<double> sv = make_view<double>(o, b, d);
        const int it = bicgstab<double, BLOCK, SWEEP>(
            sv, rhs, w.x + base, (double*)w.r + base, (double*)w.rh + base, (double*)w.p + base,
            (double*)w.v + base, (double*)w.ph + base, (double*)w.sh + base, (double*)w.t + base, cfg.maxit, cfg.tol,
            cfg.reference_semantics != 0, lds);
        if (threadIdx.x == 0) {
            w.iters[s] += (it > cfg.maxit) ? cfg.maxit : it;
            w.fails[s] += (it > cfg.maxit) ? 1 : 0;
        }
        __syncthreads();
    }
}

// Mixed precision: FP64 RHS and residual, FP32 BiCGSTAB on the scaled residual equation,
// x += ||r|| * y per refinement step, until ||b - A x|| / ||b|| < tol (FP64, on the GPU).  FP32
// BiCGSTAB on a transport operator can stall far above its target (the Jacobi-preconditioned
// operator is far from normal and FP32 rounding is amplified along long upwind chains): when an
// inner solve sets no new residual minimum for kStall32 iterations or hits the limit, or the
// refinement steps end short of tol, the work-group re-solves the system in FP64 at once (the
// FP64 solve of k_solve64), overlapping the other work-groups' FP32 solves.  The FP32 Krylov
// vectors of system s live in the first half of its FP64 slot, so that re-solve touches no other
// system's vectors.
constexpr int kStall32 = 100;

template <int BLOCK, bool SWEEP>
__global__ __launch_bounds__(BLOCK) void k_solve_mixed(Ops o, Work w, SolveCfg cfg) {
    __shared__ double lds[2 * (BLOCK / 64)];
    for (int s = blockIdx.x; s < cfg.S; s += gridDim.x) {
        const int b = s / o.D, d = s % o.D;
        const size_t base = (size_t)s * o.N;
        const double* rhs = w.b + base;
        double* x = w.x + base;
        double* r64 = w.r64 + base;
        double red[1] = {0};
        for (int c = threadIdx.x; c < o.N; c += BLOCK) {
            const double bc = rhs[c];
            r64[c] = bc;
            x[c] = 0;
            red[0] += bc * bc;
        }
        block_sum<BLOCK, 1>(red, lds);
        const double nb = sqrt(red[0]);
        double nr = nb;
        const SysView<float> sv32 = make_view<float>(o, b, d);
        const SysView<double> sv64 = make_view<double>(o, b, d);
        float* y = (float*)w.y + base;
        float* r32 = (float*)((double*)w.r + base);
        int total = 0;
        bool stalled = false;
        for (int step = 0; step < cfg.refine_steps && nr > cfg.tol * nb && nr > 1e-300; step++) {
            // y solves A y = r64 / ||r64|| in FP32 (scaled into FP32 range)
            const double inv = 1.0 / nr;
            float* fr = (float*)w.f + base;  // scaled residual as the FP32 right-hand side
            for (int c = threadIdx.x; c < o.N; c += BLOCK) fr[c] = (float)(r64[c] * inv);
            __syncthreads();
            // FP32 inner tolerance cfg.tol32 (short of FP32 rounding, where BiCGSTAB can stall);
            // the FP64 refinement supplies the rest
            const int it = bicgstab<float, BLOCK, SWEEP>(
                sv32, fr, y, r32, (float*)((double*)w.rh + base), (float*)((double*)w.p + base),
                (float*)((double*)w.v + base), (float*)((double*)w.ph + base), (float*)((double*)w.sh + base),
                (float*)((double*)w.t + base), cfg.maxit, cfg.tol32, false, lds, kStall32);
            total += (it > cfg.maxit) ? cfg.maxit : it;
            if (it > cfg.maxit) {  // stalled: y is not a correction to keep
                stalled = true;
                break;
            }
            for (int c = threadIdx.x; c < o.N; c += BLOCK) x[c] += nr * (double)y[c];
            __syncthreads();
            // r64 = b - A x in FP64
            red[0] = 0;
            for (int c = threadIdx.x; c < o.N; c += BLOCK) {
                double acc = diag_of(sv64, sv64.dg[c], c) * x[c];
                double off = 0;
                for (int k = 0; k < sv64.W; k++) {
                    const int j = sv64.col[(size_t)k * sv64.N + c];
                    if (j >= 0) off += sv64.val[(size_t)k * sv64.N + c] * x[j];
                }
                const double rc = rhs[c] - (acc + sv64.vt * off);
                r64[c] = rc;
                red[0] += rc * rc;
            }
            block_sum<BLOCK, 1>(red, lds);
            nr = sqrt(red[0]);
        }
        int fail = 0;
        if (stalled || (nr > cfg.tol * nb && nr > 1e-300)) {
            const int it = bicgstab<double, BLOCK, SWEEP>(
                sv64, rhs, x, (double*)w.r + base, (double*)w.rh + base, (double*)w.p + base, (double*)w.v + base,
                (double*)w.ph + base, (double*)w.sh + base, (double*)w.t + base, cfg.maxit, cfg.tol,
                cfg.reference_semantics != 0, lds);
            total += (it > cfg.maxit) ? cfg.maxit : it;
            fail = (it > cfg.maxit) ? 1 : 0;
        }
        if (threadIdx.x == 0) {
            w.iters[s] += total;
            w.fails[s] += fail;
        }
        __syncthreads();
    }
}

// Band temperatures (StaticBTESolver.cpp:96-104) with one thread per (band, cell) -- B times more
// threads than a cell loop, so enough loads are in flight to stream both intensity arrays at HBM
// rate -- fused with the margin partial sums (:311-321).  Work-group blk of nblk strides over the
// B*N pairs (grids are capped, B*N is not) and writes partial[blk]; each thread adds its pairs in
// increasing order, so the sums are the same on every call with the same grid.  The pairs go in
// (band, cell) order for the band-major intensity layout, in (cell, band) order (CB) for the
// interleaved one, so that neighbouring threads read neighbouring intensities either way; a
// band temperature sums its directions in the same order in both.
// U > 1: a work-group takes U*BLOCK consecutive pairs at a time and a thread U of them (BLOCK
// apart), loading all U in each direction step, so that each direction's plane is read in runs of
// U*BLOCK values; the band temperatures are the same, the margin partials add in another order.
template <int BLOCK, bool CB = false, int U = 1>
__device__ __forceinline__ void tband_block(int N, int D, const Bands& bands, const EeIdx& ei,
                                            const double* __restrict__ omega, const double* __restrict__ ee_curr,
                                            const double* __restrict__ ee_prev, double* __restrict__ Tband,
                                            double* __restrict__ partial, int blk, int nblk) {
    __shared__ double lds[BLOCK / 64];
    double mp[1] = {0};
    const long long total = (long long)bands.B * N;
    if constexpr (U > 1) {
        for (long long t0 = blk * (long long)(U * BLOCK) + threadIdx.x; t0 < total;
             t0 += (long long)nblk * (U * BLOCK)) {
            size_t base[U];
            double e0[U];
            bool on[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const long long t = t0 + (long long)u * BLOCK;
                on[u] = t < total;
                const long long tt = on[u] ? t : t0;
                const int b = CB ? (int)(tt % bands.B) : (int)(tt / N), c = CB ? (int)(tt / bands.B) : (int)(tt % N);
                base[u] = ei.at(b, 0, c);
                e0[u] = 0;
            }
#pragma unroll 2
            for (int d = 0; d < D; d++) {
                const double w = omega[d];
#pragma unroll
                for (int u = 0; u < U; u++)
                    if (on[u]) {
                        const double ec = ee_curr[base[u] + (size_t)d * ei.sd];
                        const double df = ec - ee_prev[base[u] + (size_t)d * ei.sd];
                        e0[u] += ec * w;
                        mp[0] += df * df;
                    }
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (on[u]) {
                    const long long t = t0 + (long long)u * BLOCK;
                    const int b = CB ? (int)(t % bands.B) : (int)(t / N), c = CB ? (int)(t / bands.B) : (int)(t % N);
                    Tband[(size_t)b * N + c] = e0[u] / bands.C[b] + bands.T_ref;
                }
        }
        block_sum<BLOCK, 1>(mp, lds);
        if (threadIdx.x == 0) partial[blk] = mp[0];
        return;
    }
    for (long long t = blk * (long long)BLOCK + threadIdx.x; t < total; t += (long long)nblk * BLOCK) {
        const int b = CB ? (int)(t % bands.B) : (int)(t / N), c = CB ? (int)(t / bands.B) : (int)(t % N);
        const size_t base = ei.at(b, 0, c);
        double e0 = 0;
#pragma unroll 8
        for (int d = 0; d < D; d++) {
            const double ec = ee_curr[base + (size_t)d * ei.sd];
            const double df = ec - ee_prev[base + (size_t)d * ei.sd];
            e0 += ec * omega[d];
            mp[0] += df * df;
        }
        Tband[(size_t)b * N + c] = e0 / bands.C[b] + bands.T_ref;
    }
    block_sum<BLOCK, 1>(mp, lds);
    if (threadIdx.x == 0) partial[blk] = mp[0];
}

template <int BLOCK, bool CB = false, int U = 1>
__global__ __launch_bounds__(BLOCK) void k_tband(int N, int D, Bands bands, EeIdx ei, const double* __restrict__ omega,
                                                const double* __restrict__ ee_curr, const double* __restrict__ ee_prev,
                                                double* __restrict__ Tband, double* __restrict__ partial) {
    tband_block<BLOCK, CB, U>(N, D, bands, ei, omega, ee_curr, ee_prev, Tband, partial, blockIdx.x, gridDim.x);
}

// Equilibrium temperature and emission density per cell (StaticBTESolver.cpp:106-120).
// tsrc: the temperature each cell's emission density was formed from (the source temperature)
__device__ __forceinline__ void tcell_range(int N, const Bands& bands, const double* __restrict__ Tband,
                                            double* __restrict__ T, double* __restrict__ dens,
                                            double* __restrict__ tsrc, int blk, int nblk) {
    for (int c = blk * blockDim.x + threadIdx.x; c < N; c += nblk * blockDim.x) {
        double rrn = 0, rnn = 0;
        for (int b = 0; b < bands.B; b++) {
            rrn += bands.Lr[b] * Tband[(size_t)b * N + c];
            rnn += bands.Lr[b];
        }
        const double tc = rrn / rnn;
        T[c] = tc;
        tsrc[c] = tc;
        for (int b = 0; b < bands.B; b++) dens[(size_t)c * bands.B + b] = bands.C[b] / bands.solid_angle * (tc - bands.T_ref);
    }
}

__global__ void k_tcell(int N, Bands bands, const double* __restrict__ Tband, double* __restrict__ T,
                        double* __restrict__ dens, double* __restrict__ tsrc) {
    tcell_range(N, bands, Tband, T, dens, tsrc, blockIdx.x, gridDim.x);
}

// ----------------------------------------------------------------------------------------------
// Anderson mixing of the outer iteration (fixed-point map T -> G(T) of one source iteration).
// History columns dF_j = f_{j+1} - f_j (f = G(x) - x) and dG_j = dX_j + dF_j (dX_j = x_{j+1} - x_j)
// live in ring buffers [m][Z], formed in FP64 and stored rounded to FP32: the least squares and the
// update then use columns within 6e-8 (relative) of the exact differences, which moves the mixing
// step by as little and leaves the fixed point unchanged (at g = x the step is g), and both passes
// over the history read half the bytes.  Dot products, the update and x, f, g stay FP64.  The latest
// dX alone is kept too, in FP64 (the learned correction's features).
// The m x m least-squares system is solved on the host.
// ----------------------------------------------------------------------------------------------

// partial[q * nblk + blk] for q < ncols: dF_q . dF_new, q in [ncols, 2 ncols): dF_{q-ncols} . f,
// q = 2 ncols: f . f (nblk = kAaDotBlocks)
// Grid (nblk, ceil(ncols / G) + 1): work-group (x, y < ceil(ncols / G)) reads columns
// dF_{Gy .. Gy+G-1} and, once per entry for all of them, dF_new and f; (x, last) forms f . f.  Each
// sum runs over the work-group's chunk in the same fixed order for every call, whatever G.
// (Measured on the 50-column history of a 180k-entry state, columns not in cache: 256-thread
// work-groups take 16 us with G = 2, 25 us with G = 4 and 32 us with G = 8 -- the work-group's
// cross-lane sums of 2 G values and the fewer work-groups cost more than the shared reads save.)
// Arguments of one step's dot products; out and host receive the reduced values (k_reduce_rows)
struct AaDots {
    int Z, ncols;
    const float* dF;      // the history columns, Z apart
    const float* dF_new;  // the newest column
    const double* f;
    double* partial;      // [2 ncols + 1][kAaDotBlocks]
    double* out;
    double* host;
};
constexpr int kAaDotBlocks = 64;

// Work-group (bx, by) of the (kAaDotBlocks, ceil(ncols / G) + 1) grid.
template <int BLOCK, int G>
__device__ __forceinline__ void aa_dots_block(const AaDots& a, int bx, int by) {
    __shared__ double lds[2 * G * (BLOCK / 64)];
    const int N = a.Z, ncols = a.ncols, j0 = by * G, nblk = kAaDotBlocks;
    const float* __restrict__ dF = a.dF;
    const float* __restrict__ dF_new = a.dF_new;
    const double* __restrict__ f = a.f;
    double s[2 * G];
#pragma unroll
    for (int k = 0; k < 2 * G; k++) s[k] = 0;
    if (j0 < ncols) {
        // unrolled so that several entries' loads are in flight (a thread's entries are a
        // dependent chain of sums otherwise); the sums keep their order
#pragma unroll 4
        for (int c = bx * BLOCK + threadIdx.x; c < N; c += nblk * BLOCK) {
            const double av = dF_new[c], bv = f[c];
            double v[G];
#pragma unroll
            for (int g = 0; g < G; g++) v[g] = j0 + g < ncols ? (double)dF[(size_t)(j0 + g) * N + c] : 0.0;
#pragma unroll
            for (int g = 0; g < G; g++) {
                s[g] += v[g] * av;
                s[G + g] += v[g] * bv;
            }
        }
        block_sum<BLOCK, 2 * G>(s, lds);
        if (threadIdx.x == 0)
            for (int g = 0; g < G && j0 + g < ncols; g++) {
                a.partial[(size_t)(j0 + g) * nblk + bx] = s[g];
                a.partial[(size_t)(ncols + j0 + g) * nblk + bx] = s[G + g];
            }
    } else {
#pragma unroll 4
        for (int c = bx * BLOCK + threadIdx.x; c < N; c += nblk * BLOCK) s[0] += f[c] * f[c];
        block_sum<BLOCK, 1>(s, lds);
        if (threadIdx.x == 0) a.partial[(size_t)(2 * ncols) * nblk + bx] = s[0];
    }
}

template <int BLOCK, int G>
__global__ __launch_bounds__(BLOCK) void k_aa_dots(AaDots a) {
    aa_dots_block<BLOCK, G>(a, blockIdx.x, blockIdx.y);
}

// out[q] = sum_b partial[q * nblk + b] in a fixed order; host (or null): the same values also
// stored into coherent host memory, which the host reads after synchronising (no copy)
__device__ __forceinline__ void reduce_rows_range(int nq, int nblk, const double* __restrict__ partial,
                                                  double* __restrict__ out, double* __restrict__ host, int blk,
                                                  int ngrid) {
    __shared__ double lds[4];
    for (int q = blk; q < nq; q += ngrid) {
        double s[1] = {0};
        for (int i = threadIdx.x; i < nblk; i += 256) s[0] += partial[(size_t)q * nblk + i];
        block_sum<256, 1>(s, lds);
        if (threadIdx.x == 0) {
            out[q] = s[0];
            if (host) host[q] = s[0];
        }
    }
}

__global__ void k_reduce_rows(int nq, int nblk, const double* __restrict__ partial, double* __restrict__ out,
                              double* __restrict__ host) {
    reduce_rows_range(nq, nblk, partial, out, host, blockIdx.x, gridDim.x);
}

// Stores seq into coherent host memory; launched behind the work a host thread waits for, it runs
// once all of it has completed, so the host can poll that word instead of synchronising the stream.
__global__ void k_signal(long long seq, long long* __restrict__ host) {
    if (threadIdx.x == 0) __hip_atomic_store(host, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// The state of the outer iteration is the temperature field plus the intensities at cells with a
// reflecting (diffuse or specular) boundary face -- the only parts of the previous iterate the next
// right-hand sides read.  Intensities are carried as temperature equivalents e * Omega / C_b.
// z = [T (N) ; e_bnd (S x nb)].
// g = G(x) gathered (T, and the boundary intensities at bcells scaled by sc), and in the same pass
// f = g - x into f_cur and f_prev, x into x_prev and, unless first, the history columns
// dX = x - x_prev and dF = f - f_prev (from the previous step's values), over the Z = N + S nb
// state entries.  take: then x = g (a plain step), and with dens the state's loads for the next
// iteration from it, as k_set_density and k_state_scatter form them (the emission density of the T
// part, the boundary intensities).
struct AaGather {
    int N, D, S, nb;
    EeIdx ei;
    const int* bcells;
    const double* sc;
    const double* T;
    double* ee;
    double* z;
    int first, take;
    double* x;
    double* x_prev;
    double* f_prev;
    double* f_cur;
    double* dX;
    float* dF;
    float* dG;
    Bands bands;
    double* dens;
    double* tsrc;
};

// Body over work-group blk of nblk (blockDim.x threads each).  Every load of an entry precedes its
// stores.
__device__ __forceinline__ void gather_diff_range(const AaGather& a, int blk, int nblk) {
    const int N = a.N, D = a.D, nb = a.nb;
    const long long total = N + (long long)a.S * nb;
    for (long long t = blk * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)nblk * blockDim.x) {
        double gv;
        long long at = 0;  // the boundary intensity's index
        int s = 0;
        if (t < N) {
            gv = a.T[t];
        } else {
            const long long u = t - N;
            s = (int)(u / nb);
            const int j = (int)(u % nb);
            at = (long long)a.ei.at(s / D, s % D, a.bcells[j]);
            gv = a.ee[at] * a.sc[s / D];
        }
        const double xv = a.x[t];
        const double f = gv - xv;
        double dx = 0, df = 0;
        if (!a.first) {
            dx = xv - a.x_prev[t];
            df = f - a.f_prev[t];
        }
        a.z[t] = gv;
        if (!a.first) {
            a.dX[t] = dx;
            a.dF[t] = (float)df;
            a.dG[t] = (float)(dx + df);
        }
        a.x_prev[t] = xv;
        a.f_prev[t] = f;
        a.f_cur[t] = f;
        if (a.take) {
            a.x[t] = gv;
            if (a.dens) {
                if (t < N) {
                    a.tsrc[t] = gv;
                    for (int b = 0; b < a.bands.B; b++)
                        a.dens[(size_t)t * a.bands.B + b] = a.bands.C[b] / a.bands.solid_angle * (gv - a.bands.T_ref);
                } else {
                    a.ee[at] = gv / a.sc[s / D];
                }
            }
        }
    }
}

__global__ void k_state_gather_diff(AaGather a) { gather_diff_range(a, blockIdx.x, gridDim.x); }

__global__ void k_state_scatter(int N, int D, int S, int nb, EeIdx ei, const int* __restrict__ bcells,
                                const double* __restrict__ sc, const double* __restrict__ z, double* __restrict__ ee) {
    const long long total = (long long)S * nb;
    for (long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x; u < total;
         u += (long long)gridDim.x * blockDim.x) {
        const int s = (int)(u / nb), j = (int)(u % nb);
        ee[ei.at(s / D, s % D, bcells[j])] = z[N + u] / sc[s / D];
    }
}

// ----------------------------------------------------------------------------------------------
// Learned correction.  Node features of the state the next outer iteration would start from
// (temperature part of the fixed-point state; x = last evaluated state, f = G(x) - x its
// residual, x_next = the mixed next state, s = RMS of f over cells):
//   0 f / s      1 (x_next - x) / s      2 (g - g_prev) / s (last history column, 0 without one)
//   3 (x_next - T_mid) / T_span      4 log10(s / T_span) / 4 (uniform)
//   5 (x_next - T_wall_interp) / T_span
//   6..15 static per-cell features (wall flags, graph distances, Knudsen numbers; host-built)
// The network output y is the correction T_span * y to the temperature part of x_next.
// ----------------------------------------------------------------------------------------------
constexpr int kLearnDynamic = 6;
constexpr int kLearnStatic = 16 - kLearnDynamic;

// ffT: device address of sum_c f_c^2 over the temperature part (k_sqdiff2), from which s follows.
__global__ void k_learn_features(int N, const double* __restrict__ f, const double* __restrict__ x_next,
                                 const double* __restrict__ x, const double* __restrict__ dXl,
                                 const float* __restrict__ dFl, const float* __restrict__ stat,
                                 const double* __restrict__ ffT, double T_mid, double span, float* __restrict__ X) {
    const double sv = sqrt(*ffT / N), inv_span = 1.0 / span;
    const double inv_s = sv > 0 ? 1.0 / sv : 0.0;
    const float log_s = sv > 0 ? (float)(log10(sv / span) / 4) : -2.f;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
        float* o = X + (size_t)c * 16;
        const float* st = stat + (size_t)c * kLearnStatic;
        const float t = (float)((x_next[c] - T_mid) * inv_span);
        o[0] = (float)(f[c] * inv_s);
        o[1] = (float)((x_next[c] - x[c]) * inv_s);
        o[2] = dXl ? (float)((dXl[c] + dFl[c]) * inv_s) : 0.f;
        o[3] = t;
        o[4] = log_s;
        o[5] = t - st[3];
#pragma unroll
        for (int k = 0; k < kLearnStatic; k++) o[kLearnDynamic + k] = st[k];
    }
}

// z += s * y on the temperature part and, as a local equilibrium shift, on the reflecting-boundary
// intensities of each corrected cell (they are carried as temperature equivalents).
// The learned correction z += span y (per cell; a boundary entry takes its cell's), with the state
// before it kept in backup and the state's loads for the next iteration formed from the corrected
// z as k_set_density and k_state_scatter form them.  ff: device address of the proposal's residual
// norms {over T, over the whole state}; a state with no residual (or a non-finite one) is only
// backed up.
__global__ void k_apply_correction(int N, int D, int S, int nb, EeIdx ei, const int* __restrict__ bcells,
                                   const float* __restrict__ y, double span, const double* __restrict__ ff,
                                   double* __restrict__ z, double* __restrict__ backup, Bands bands,
                                   double* __restrict__ dens, double* __restrict__ tsrc, const double* __restrict__ sc,
                                   double* __restrict__ ee) {
    const bool apply = ff[0] > 0 && isfinite(ff[1]);
    const long long total = N + (long long)S * nb;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const double v0 = z[t];
        backup[t] = v0;
        if (!apply) continue;
        const int c = t < N ? (int)t : bcells[(t - N) % nb];
        const double v = v0 + span * (double)y[c];
        z[t] = v;
        if (t < N) {
            tsrc[t] = v;
            for (int b = 0; b < bands.B; b++)
                dens[(size_t)t * bands.B + b] = bands.C[b] / bands.solid_angle * (v - bands.T_ref);
        } else {
            const long long u = t - N;
            const int s = (int)(u / nb), j = (int)(u % nb);
            ee[ei.at(s / D, s % D, bcells[j])] = v / sc[s / D];
        }
    }
}

// partial sums of sum (a - b)^2 over [lo, hi)
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sqdiff(long long lo, long long hi, const double* __restrict__ a,
                                                 const double* __restrict__ b, double* __restrict__ partial) {
    __shared__ double lds[BLOCK / 64];
    double s[1] = {0};
    for (long long i = lo + blockIdx.x * (long long)BLOCK + threadIdx.x; i < hi; i += (long long)gridDim.x * BLOCK) {
        const double d = a[i] - (b ? b[i] : 0.0);
        s[0] += d * d;
    }
    block_sum<BLOCK, 1>(s, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// Two squared norms in one pass: over [0, n1) of (a1 - b1)^2 and over [0, n2) of (a2 - b2)^2 (a null
// b counts as 0); one partial sum per work-group and norm, partial[q * gridDim.x + blk], each summed
// in the same fixed order for every call.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_sqdiff2(long long n1, const double* __restrict__ a1,
                                                  const double* __restrict__ b1, long long n2,
                                                  const double* __restrict__ a2, const double* __restrict__ b2,
                                                  double* __restrict__ partial) {
    __shared__ double lds[2 * (BLOCK / 64)];
    double s[2] = {0, 0};
    const long long n = n1 > n2 ? n1 : n2, st = (long long)gridDim.x * BLOCK;
    for (long long i = blockIdx.x * (long long)BLOCK + threadIdx.x; i < n; i += st) {
        if (i < n1) {
            const double d = a1[i] - (b1 ? b1[i] : 0.0);
            s[0] += d * d;
        }
        if (i < n2) {
            const double d = a2[i] - (b2 ? b2[i] : 0.0);
            s[1] += d * d;
        }
    }
    block_sum<BLOCK, 2>(s, lds);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = s[0];
        partial[gridDim.x + blockIdx.x] = s[1];
    }
}

// The judge's two squared norms in one pass, summed as k_sqdiff2 sums them: over the Z = N + S nb
// state entries of (g - x) with g = G(x~) gathered as k_state_gather_diff does (not stored), and over
// [0, N) of (x - backup).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_judge_norms(int N, int D, int S, int nb, EeIdx ei,
                                                      const int* __restrict__ bcells,
                                                      const double* __restrict__ sc, const double* __restrict__ T,
                                                      const double* __restrict__ ee, const double* __restrict__ x,
                                                      const double* __restrict__ backup, double* __restrict__ partial) {
    __shared__ double lds[2 * (BLOCK / 64)];
    double acc[2] = {0, 0};
    const long long Z = N + (long long)S * nb, st = (long long)gridDim.x * BLOCK;
    for (long long i = blockIdx.x * (long long)BLOCK + threadIdx.x; i < Z; i += st) {
        double gv;
        if (i < N) {
            gv = T[i];
        } else {
            const long long u = i - N;
            const int s = (int)(u / nb), j = (int)(u % nb);
            gv = ee[ei.at(s / D, s % D, bcells[j])] * sc[s / D];
        }
        const double d = gv - x[i];
        acc[0] += d * d;
        if (i < N) {
            const double e = x[i] - backup[i];
            acc[1] += e * e;
        }
    }
    block_sum<BLOCK, 2>(acc, lds);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = acc[0];
        partial[gridDim.x + blockIdx.x] = acc[1];
    }
}

__global__ void k_fill(int n, double v, double* __restrict__ x) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = v;
}

// Mixing coefficients of one Anderson step.
constexpr int kAaMaxCols = 64;
struct AaCoef {
    double g[kAaMaxCols];
    __device__ double operator[](int j) const { return g[j]; }
};

struct AaUpdate {
    int N, D, S, nb;
    EeIdx ei;
    int ncols;
    const double* g;
    const float* dG;
    double* x;
    Bands bands;
    double* dens;
    const int* bcells;
    const double* sc;
    double* ee;
    const double* gamma;  // the coefficients in device memory (the multi-problem launch)
    double* tsrc;
};

// The mixing update x = g - sum_j gamma_j (dX_j + dF_j) over the Z = N + S nb state entries (the
// columns dG_j = dX_j + dF_j are stored as such), and in the same pass the state's loads for the
// next iteration from the new x: the emission density of its temperature part (as k_set_density)
// and its boundary intensities (as k_state_scatter).  Body over work-group blk of nblk.
template <class Coef>
__device__ __forceinline__ void aa_update_range(const AaUpdate& a, const Coef& gamma, int blk, int nblk) {
    const long long Z = a.N + (long long)a.S * a.nb;
    const float* __restrict__ dG = a.dG;
    for (long long c = blk * (long long)blockDim.x + threadIdx.x; c < Z; c += (long long)nblk * blockDim.x) {
        double v = a.g[c];
        // unrolled so that several columns' loads are in flight; the updates keep their order
#pragma unroll 16
        for (int j = 0; j < a.ncols; j++) v -= gamma[j] * (double)dG[(size_t)j * Z + c];
        a.x[c] = v;
        if (c < a.N) {
            a.tsrc[c] = v;
            for (int b = 0; b < a.bands.B; b++)
                a.dens[(size_t)c * a.bands.B + b] = a.bands.C[b] / a.bands.solid_angle * (v - a.bands.T_ref);
        } else {
            const long long u = c - a.N;
            const int s = (int)(u / a.nb), j = (int)(u % a.nb);
            a.ee[a.ei.at(s / a.D, s % a.D, a.bcells[j])] = v / a.sc[s / a.D];
        }
    }
}

// Single problem: the coefficients travel as kernel arguments (no upload precedes the update).
__global__ void k_aa_update_state(AaUpdate a, AaCoef gamma) { aa_update_range(a, gamma, blockIdx.x, gridDim.x); }

// Emission density of a given temperature field (StaticBTESolver.cpp:115-119), cell-major [N][B].
__global__ void k_set_density(int N, Bands bands, const double* __restrict__ T, double* __restrict__ dens,
                              double* __restrict__ tsrc) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
        tsrc[c] = T[c];
        for (int b = 0; b < bands.B; b++)
            dens[(size_t)c * bands.B + b] = bands.C[b] / bands.solid_angle * (T[c] - bands.T_ref);
    }
}

// Sum of the block partials in a fixed order; margin = sum / (N_band * N_cell).  host (or null):
// the margin also stored into coherent host memory, which the host reads after synchronising.
__device__ __forceinline__ void margin_block(int nparts, const double* __restrict__ partial, double scale,
                                             double* __restrict__ out, double* __restrict__ host = nullptr) {
    __shared__ double lds[4];
    double s[1] = {0};
    for (int i = threadIdx.x; i < nparts; i += 256) s[0] += partial[i];
    block_sum<256, 1>(s, lds);
    if (threadIdx.x == 0) {
        out[0] = s[0] * scale;
        if (host) host[0] = s[0] * scale;
    }
}

__global__ void k_margin(int nparts, const double* __restrict__ partial, double scale, double* __restrict__ out) {
    margin_block(nparts, partial, scale, out);
}

// k_tcell over work-groups [0, n_tcell) and k_margin as work-group n_tcell, in one launch: both
// read only the band-temperature step's results.
__global__ __launch_bounds__(256) void k_tcell_margin(int N, Bands bands, const double* __restrict__ Tband,
                                                      double* __restrict__ T, double* __restrict__ dens,
                                                      double* __restrict__ tsrc, int n_tcell, int nparts,
                                                      const double* __restrict__ partial, double scale,
                                                      double* __restrict__ out, double* __restrict__ host) {
    if ((int)blockIdx.x < n_tcell)
        tcell_range(N, bands, Tband, T, dens, tsrc, blockIdx.x, n_tcell);
    else
        margin_block(nparts, partial, scale, out, host);
}

// ----------------------------------------------------------------------------------------------
// Uneven batch: the per-problem steps of an outer iteration around the batched sweep, one launch
// per step for all problems.  Work-group t belongs to the problem p with off[p] <= t < off[p + 1]
// and runs that problem's body as work-group t - off[p] of off[p + 1] - off[p], with the same
// partition (and so the same reduction order) as the single-problem launch.
// ----------------------------------------------------------------------------------------------
struct IterDesc {
    Geo g;
    Quad q;
    Bnd bd;
    Bands bands;
    Ops o;
    const double* ee_prev;  // previous iterate
    const double* ee_curr;  // the sweep's new iterate
    double* ediff;          // diffuse re-emission, written by the diffuse step
    double* rhs;            // level-order right-hand sides
    const int* bnd_list;    // level-order entries with an incoming boundary face
    int n_bnd_list;
    double* Tband;
    double* T;
    double* dens;
    double* tsrc;
    const double* tsrc_rhs;  // tsrc when it holds (the right-hand sides are formed from it), else null
    double* partial;
    double* margin;
    const double* omega;
    int nblk_tband;         // work-groups of the band-temperature step (its margin partials)
    double scale;           // 1 / (B N)
};

struct MultiIdx {
    int p, blk, nblk;
};

__device__ __forceinline__ MultiIdx multi_index(const int* __restrict__ off, int n) {
    const int t = blockIdx.x;
    int lo = 0, hi = n;  // the last p with off[p] <= t
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= t)
            lo = mid;
        else
            hi = mid;
    }
    return MultiIdx{lo, t - off[lo], off[lo + 1] - off[lo]};
}

__global__ void k_diffuse_multi(cmem<IterDesc>* __restrict__ desc, const int* __restrict__ off, int n) {
    const MultiIdx m = multi_index(off, n);
    cmem<IterDesc>* P = desc + m.p;
    const Geo g = cload(&P->g);
    const Quad q = cload(&P->q);
    diffuse_range(g, q, bnd_view(&P->bd), P->bands.B, P->ee_prev, P->ediff, m.blk, m.nblk);
}

__global__ void k_rhs_levels_multi(cmem<IterDesc>* __restrict__ desc, const int* __restrict__ off, int n) {
    const MultiIdx m = multi_index(off, n);
    const IterDesc D = cload(desc + m.p);
    rhs_levels_range(D.o, D.bands.B, D.dens, D.tsrc_rhs, D.rhs, m.blk, m.nblk);
}

__global__ void k_rhs_bnd_multi(cmem<IterDesc>* __restrict__ desc, const int* __restrict__ off, int n) {
    const MultiIdx m = multi_index(off, n);
    cmem<IterDesc>* P = desc + m.p;
    const Geo g = cload(&P->g);
    const Quad q = cload(&P->q);
    const Bands bands = cload(&P->bands);
    const Ops o = cload(&P->o);
    rhs_bnd_range(g, q, bnd_view(&P->bd), bands, o, bands.B * o.D, P->bnd_list, P->n_bnd_list, P->ee_prev, P->dens,
                  P->rhs, m.blk, m.nblk);
}

__global__ __launch_bounds__(256) void k_tband_multi(cmem<IterDesc>* __restrict__ desc, const int* __restrict__ off,
                                                     int n) {
    const MultiIdx m = multi_index(off, n);
    const IterDesc D = cload(desc + m.p);
    tband_block<256>(D.o.N, D.o.D, D.bands, D.q.ee, D.omega, D.ee_curr, D.ee_prev, D.Tband, D.partial, m.blk, m.nblk);
}

__global__ void k_tcell_multi(cmem<IterDesc>* __restrict__ desc, const int* __restrict__ off, int n) {
    const MultiIdx m = multi_index(off, n);
    const IterDesc D = cload(desc + m.p);
    tcell_range(D.o.N, D.bands, D.Tband, D.T, D.dens, D.tsrc, m.blk, m.nblk);
}

// one work-group per problem
__global__ __launch_bounds__(256) void k_margin_multi(cmem<IterDesc>* __restrict__ desc) {
    const IterDesc D = cload(desc + blockIdx.x);
    margin_block(D.nblk_tband, D.partial, D.scale, D.margin);
}

// The mixing steps of the problems, one launch per kernel (AaGather, AaDots, AaUpdate descriptors);
// each problem's work-groups keep the single-problem launch's partition: grid_for(Z, 256) for the
// gather and the update, (kAaDotBlocks, ceil(ncols / G) + 1) for the dot products (flattened
// x-fastest), min(2 ncols + 1, 128) for their reduction.
__global__ void k_state_gather_diff_multi(cmem<AaGather>* __restrict__ desc, const int* __restrict__ off, int n) {
    const MultiIdx m = multi_index(off, n);
    const AaGather D = cload(desc + m.p);
    gather_diff_range(D, m.blk, m.nblk);
}

template <int BLOCK, int G>
__global__ __launch_bounds__(BLOCK) void k_aa_dots_multi(cmem<AaDots>* __restrict__ desc, const int* __restrict__ off,
                                                        int n) {
    const MultiIdx m = multi_index(off, n);
    const AaDots D = cload(desc + m.p);
    aa_dots_block<BLOCK, G>(D, m.blk % kAaDotBlocks, m.blk / kAaDotBlocks);
}

__global__ void k_aa_reduce_multi(cmem<AaDots>* __restrict__ desc, const int* __restrict__ off, int n) {
    const MultiIdx m = multi_index(off, n);
    const AaDots D = cload(desc + m.p);
    reduce_rows_range(2 * D.ncols + 1, kAaDotBlocks, D.partial, D.out, D.host, m.blk, m.nblk);
}

__global__ void k_aa_update_multi(cmem<AaUpdate>* __restrict__ desc, const int* __restrict__ off, int n) {
    const MultiIdx m = multi_index(off, n);
    const AaUpdate D = cload(desc + m.p);
    aa_update_range(D, D.gamma, m.blk, m.nblk);
}

// Heat flux per (BC, band), StaticBTESolver.cpp:323-370.  One work-group per pair.
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_flux(Geo g, Quad q, Bnd bd, Bands bands, int DM,
                                               const double* __restrict__ ee, double* __restrict__ band_flux) {
    __shared__ double lds[2 * (BLOCK / 64)];
    const int k = blockIdx.x / bands.B, b = blockIdx.x % bands.B;
    double red[2] = {0, 0};
    for (int bf = threadIdx.x; bf < bd.n_bf; bf += BLOCK) {
        const int c = bd.bf_cell[bf], f = bd.bf_face[bf];
        if (g.nbr[(size_t)c * g.F + f] != bd.bci[k]) continue;
        const double* n = g.normal + 3 * ((size_t)c * g.F + f);
        const double A = g.area[(size_t)c * g.F + f];
        double heat = 0;
        for (int d = 0; d < q.D; d++) {
            const double e = ee[q.ee.at(b, d, c)];
            if (DM == 3 && g.dim == 3)
                heat += e * dot_nc(q.dir + 3 * d, n) * A * q.omega[d];
            else if (g.dim == 1)
                heat += e * dot_nc(q.dir + 3 * d, n) * q.omega[d];
            else
                heat += e * dot_nc(q.S + 3 * d, n) * A;
        }
        red[0] += heat;
        red[1] += A;
    }
    block_sum<BLOCK, 2>(red, lds);
    if (threadIdx.x == 0) {
        const double qv = (g.dim != 1) ? red[0] / red[1] * bands.v[b] : red[0] * bands.v[b];
        band_flux[(size_t)k * bands.B + b] = qv;
    }
}

}  // namespace bte
