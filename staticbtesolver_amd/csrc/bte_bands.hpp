// Band-batched BiCGSTAB (Jacobi preconditioner): one work-group solves the systems of G bands of
// one direction together.
//
// The operator of a direction is the same for every band up to the scalar v_b tau_b of the face
// terms and the band factor of the diagonal (v_b tau_b * face sum + w_d V_c), so every pass loads
// a cell's operator row (ELL columns and values, diagonal face sum, volume) once and applies it to
// the G bands' vectors.  Each band keeps its own scalars and stopping test; a band that has
// converged (or does not exist in a short last group) is masked out of every later pass, so its
// vectors are no longer streamed.  Per band, the arithmetic and the order of every per-thread sum
// are those of bicgstab() with the Jacobi preconditioner (bicgstab.hpp:146-298 semantics with
// reference_semantics, the standard form otherwise).
//
// Addressing: every array is read and written with buffer instructions.  A buffer resource (four
// scalar registers) holds the base of the array slice of the group's first band; band g is
// reached through a scalar byte offset g * D * N * sizeof(R), and a lane's cell through one
// vector byte offset c * sizeof(R) shared by every array and band.  With plain pointers each
// (array, band) pair costs a 64-bit vector address per lane.
#pragma once

#include "bte_kernels.hpp"

namespace bte {

typedef unsigned int bte_u32x2 __attribute__((ext_vector_type(2)));

// Raw buffer resource over [p, p + bytes) (stride 0; word 3 as for every gfx9 target).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void* p, size_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)(unsigned)bytes, 0x00020000);
}

template <typename T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

template <typename T>
__device__ __forceinline__ void buf_st(T v, __amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(bte_u32x2, v), r, voff, soff, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, voff, soff, 0);
}

// One [B*D*N] vector seen from a band group: element c of band g.
template <typename R>
struct GVec {
    __amdgpu_buffer_rsrc_t rs;
    __device__ __forceinline__ R ld(int c, unsigned so) const { return buf_ld<R>(rs, (unsigned)c * sizeof(R), so); }
    __device__ __forceinline__ void st(int c, unsigned so, R v) const {
        buf_st<R>(v, rs, (unsigned)c * sizeof(R), so);
    }
};

template <typename R, int G>
struct BandGroup {
    int N, W;
    R wv;                   // w_d
    __amdgpu_buffer_rsrc_t dg, vol, col, val;  // the direction's operator rows
    R vt[G];                // v_b tau_b of each band
    unsigned so[G];         // byte offset of band g's system from the group's first: g D N sizeof(R)
    int sys[G];             // system index s = b*D + d
    unsigned present;       // bands of the group that exist
    size_t base0;           // element offset of the first band's system
    size_t span;            // elements from base0 to the end of the last present band's system
};

template <typename R, int G>
__device__ __forceinline__ BandGroup<R, G> make_group(const Ops& o, int B, int d, int b0) {
    BandGroup<R, G> bg;
    const size_t N = (size_t)o.N;
    bg.N = o.N;
    bg.W = o.W;
    bg.wv = (R)o.omega[d];
    const R* dg;
    const R* val;
    if constexpr (sizeof(R) == 8) {
        dg = (const R*)o.dg;
        val = (const R*)o.ell_val;
    } else {
        dg = (const R*)o.dg32;
        val = (const R*)o.ell_val32;
    }
    bg.dg = buffer_of(dg + d * N, N * sizeof(R));
    bg.vol = buffer_of(o.vol, N * sizeof(double));
    bg.col = buffer_of(o.ell_col + (size_t)d * o.W * N, (size_t)o.W * N * sizeof(int));
    bg.val = buffer_of(val + (size_t)d * o.W * N, (size_t)o.W * N * sizeof(R));
    bg.present = 0;
    int last = 0;
#pragma unroll
    for (int g = 0; g < G; g++) {
        const int b = b0 + g;
        const bool ok = b < B;
        bg.vt[g] = ok ? (R)o.vt[b] : (R)0;
        bg.sys[g] = ok ? b * o.D + d : 0;
        bg.so[g] = (unsigned)((size_t)g * o.D * N * sizeof(R));
        if (ok) {
            bg.present |= 1u << g;
            last = g;
        }
    }
    bg.base0 = ((size_t)b0 * o.D + d) * N;
    bg.span = (size_t)last * o.D * N + N;
    return bg;
}

template <typename R, int G>
__device__ __forceinline__ GVec<R> group_vec(const BandGroup<R, G>& bg, const void* p) {
    return GVec<R>{buffer_of((const R*)p + bg.base0, bg.span * sizeof(R))};
}

template <int G, class F>
__device__ __forceinline__ void each_band(unsigned mask, F&& f) {
#pragma unroll
    for (int g = 0; g < G; g++)
        if (mask & (1u << g)) f(g);
}

template <typename R>
struct KrylovVecs {
    GVec<R> rhs, x, r, rh, p, v, ph, sh, t;
};

// Off-diagonal sums of the G bands' products at cell c: off[g] = sum_f a_f x_g[j_f], the row
// loaded once.  Slots beyond the width and missing neighbours (column -1) contribute nothing; the
// gather of a missing neighbour reads the cell itself, so no load is conditional.
template <typename R, int G>
__device__ __forceinline__ void group_offdiag(const BandGroup<R, G>& bg, unsigned m, const GVec<R>& x, int c,
                                              R (&off)[G]) {
    const int W = bg.W;
    int j[3];
    R a[3];
#pragma unroll
    for (int f = 0; f < 3; f++) {
        const unsigned fo = (unsigned)f * (unsigned)bg.N;
        j[f] = (f < W) ? buf_ld<int>(bg.col, (unsigned)c * 4u, fo * 4u) : -1;
        a[f] = (f < W) ? buf_ld<R>(bg.val, (unsigned)c * sizeof(R), fo * sizeof(R)) : (R)0;
    }
    R gx[G][3];
    each_band<G>(m, [&](int g) {
#pragma unroll
        for (int f = 0; f < 3; f++) gx[g][f] = x.ld(j[f] >= 0 ? j[f] : c, bg.so[g]);
    });
    each_band<G>(m, [&](int g) {
        R o = 0;
#pragma unroll
        for (int f = 0; f < 3; f++)
            if (j[f] >= 0) o += a[f] * gx[g][f];
        off[g] = o;
    });
}

template <typename R, int G>
__device__ __forceinline__ R group_diag(const BandGroup<R, G>& bg, int g, R dgc, double vl) {
    return bg.vt[g] * dgc + bg.wv * (R)vl;
}

// Per band, its[g] = iterations used (maxit + 1: limit hit) for the bands in act.
//
// The per-band scalars are the same in every lane; they live in LDS (st, 7 G doubles) instead of
// taking registers in every lane.  Every lane writes the same values, each variable right after the
// reduction that produces it; its previous value was last read before at least one barrier that
// precedes every write, except for rho, which is read (as rho_{i-1}) and written in the same phase
// and therefore alternates between two slots by iteration parity.
template <typename R, int BLOCK, int G>
__device__ void bicgstab_bands(const BandGroup<R, G>& bg, unsigned act, const KrylovVecs<R>& V, int maxit,
                               double tol, bool refsem, double* lds, double* st, int (&its)[G]) {
    const int N = bg.N;
    double* const s_rho = st;            // [2][G] by iteration parity
    double* const s_alpha = st + 2 * G;  // [G]
    double* const s_omega = st + 3 * G;
    double* const s_rnext = st + 4 * G;
    double* const s_nrm0 = st + 5 * G;
    double* const s_nrm = st + 6 * G;
    double red[2 * G];
    // setup: x = 0, r = rhs, r_hat = r, p = r (reference) or 0, v = 0
#pragma unroll
    for (int g = 0; g < G; g++) red[g] = 0;
    for (int c = threadIdx.x; c < N; c += BLOCK) {
        R bc[G];
        each_band<G>(act, [&](int g) { bc[g] = V.rhs.ld(c, bg.so[g]); });
        each_band<G>(act, [&](int g) {
            const unsigned so = bg.so[g];
            V.x.st(c, so, (R)0);
            V.r.st(c, so, bc[g]);
            V.rh.st(c, so, bc[g]);
            V.p.st(c, so, refsem ? bc[g] : (R)0);
            V.v.st(c, so, (R)0);
            red[g] += (double)bc[g] * (double)bc[g];
        });
    }
    block_sum<BLOCK, G>(red, lds);
#pragma unroll
    for (int g = 0; g < G; g++) {
        s_nrm0[g] = sqrt(red[g]);
        s_nrm[g] = sqrt(red[g]);
        s_rho[g] = s_rho[G + g] = s_alpha[g] = s_omega[g] = 1;
        s_rnext[g] = red[g];
        its[g] = maxit + 1;
    }
    // breakdown restart of the bands in m (bicgstab.hpp:170-180, :224-235); cur = rho's slot
    auto restart = [&](unsigned m, int cur) {
        for (int c = threadIdx.x; c < N; c += BLOCK) {
            R bc[G];
            each_band<G>(m, [&](int g) { bc[g] = V.rhs.ld(c, bg.so[g]); });
            each_band<G>(m, [&](int g) {
                const unsigned so = bg.so[g];
                V.x.st(c, so, (R)0);
                V.v.st(c, so, (R)0);
                V.p.st(c, so, (R)0);
                V.r.st(c, so, bc[g]);
                V.rh.st(c, so, refsem ? (R)(bc[g] * (R)6.8 + (R)0.6) : bc[g]);
            });
        }
        __syncthreads();
#pragma unroll
        for (int g = 0; g < G; g++) red[g] = 0;
        for (int c = threadIdx.x; c < N; c += BLOCK)
            each_band<G>(m, [&](int g) {
                red[g] += (double)V.rh.ld(c, bg.so[g]) * (double)V.r.ld(c, bg.so[g]);
            });
        block_sum<BLOCK, G>(red, lds);
        each_band<G>(m, [&](int g) {
            s_rho[cur * G + g] = s_alpha[g] = s_omega[g] = 1;
            s_rnext[g] = red[g];
        });
    };
    for (int i = 1; i <= maxit && act; i++) {
        const int cur = i & 1, prev = cur ^ 1;
        unsigned wm = 0, rs = 0;
        R om[G], be[G];
        each_band<G>(act, [&](int g) {
            if (s_nrm[g] < 1e-50) {
                its[g] = i;
                act &= ~(1u << g);
                return;
            }
            const double rhop = s_rho[prev * G + g], rho = s_rnext[g];
            s_rho[cur * G + g] = rho;
            if (rho < 1e-200 && rho > -1e-200) {
                rs |= 1u << g;
            } else {
                wm |= 1u << g;
                om[g] = (R)s_omega[g];
                be[g] = (R)((rho / rhop) * (s_alpha[g] / s_omega[g]));
            }
        });
        if (rs) restart(rs, cur);
        if (!wm) continue;
        // (1) p = (p - omega v) * beta + r, p^ = p / diag
        for (int c = threadIdx.x; c < N; c += BLOCK) {
            const R dgc = buf_ld<R>(bg.dg, (unsigned)c * sizeof(R), 0);
            const double vl = buf_ld<double>(bg.vol, (unsigned)c * 8u, 0);
            R pv[G], vv[G], rv[G];
            each_band<G>(wm, [&](int g) {
                pv[g] = V.p.ld(c, bg.so[g]);
                vv[g] = V.v.ld(c, bg.so[g]);
                rv[g] = V.r.ld(c, bg.so[g]);
            });
            each_band<G>(wm, [&](int g) {
                const R pn = (pv[g] - om[g] * vv[g]) * be[g] + rv[g];
                V.p.st(c, bg.so[g], pn);
                V.ph.st(c, bg.so[g], pn / group_diag(bg, g, dgc, vl));
            });
        }
        __syncthreads();
        // (2) v = A p^ (diagonal term p), <r_hat, v>
#pragma unroll
        for (int g = 0; g < G; g++) red[g] = 0;
        for (int c = threadIdx.x; c < N; c += BLOCK) {
            R dz[G], rhv[G], off[G];
            each_band<G>(wm, [&](int g) {
                dz[g] = V.p.ld(c, bg.so[g]);
                rhv[g] = V.rh.ld(c, bg.so[g]);
            });
            group_offdiag<R, G>(bg, wm, V.ph, c, off);
            each_band<G>(wm, [&](int g) {
                const R vc = dz[g] + bg.vt[g] * off[g];
                V.v.st(c, bg.so[g], vc);
                red[g] += (double)rhv[g] * (double)vc;
            });
        }
        block_sum<BLOCK, G>(red, lds);
        rs = 0;
        each_band<G>(wm, [&](int g) {
            if (red[g] < 1e-200 && red[g] > -1e-200) {
                rs |= 1u << g;
                wm &= ~(1u << g);
            } else {
                s_alpha[g] = s_rho[cur * G + g] / red[g];
            }
        });
        if (rs) restart(rs, cur);
        if (!wm) continue;
        // (3) s = r - alpha v (in r), ||s||^2, s^ = s / diag
        {
            R al[G];
            each_band<G>(wm, [&](int g) { al[g] = (R)s_alpha[g]; });
#pragma unroll
            for (int g = 0; g < G; g++) red[g] = 0;
            for (int c = threadIdx.x; c < N; c += BLOCK) {
                const R dgc = buf_ld<R>(bg.dg, (unsigned)c * sizeof(R), 0);
                const double vl = buf_ld<double>(bg.vol, (unsigned)c * 8u, 0);
                R rv[G], vv[G];
                each_band<G>(wm, [&](int g) {
                    rv[g] = V.r.ld(c, bg.so[g]);
                    vv[g] = V.v.ld(c, bg.so[g]);
                });
                each_band<G>(wm, [&](int g) {
                    const R s = rv[g] - al[g] * vv[g];
                    V.r.st(c, bg.so[g], s);
                    V.sh.st(c, bg.so[g], s / group_diag(bg, g, dgc, vl));
                    red[g] += (double)s * (double)s;
                });
            }
        }
        block_sum<BLOCK, G>(red, lds);
        if (!refsem) {
            unsigned ex = 0;
            each_band<G>(wm, [&](int g) {
                if (sqrt(red[g]) < tol * s_nrm0[g]) ex |= 1u << g;
            });
            if (ex) {
                R al[G];
                each_band<G>(ex, [&](int g) { al[g] = (R)s_alpha[g]; });
                for (int c = threadIdx.x; c < N; c += BLOCK)
                    each_band<G>(ex, [&](int g) {
                        V.x.st(c, bg.so[g], V.x.ld(c, bg.so[g]) + al[g] * V.ph.ld(c, bg.so[g]));
                    });
                each_band<G>(ex, [&](int g) { its[g] = i; });
                act &= ~ex;
                wm &= ~ex;
                __syncthreads();
                if (!wm) continue;
            }
        }
        // (4) t = A s^ (diagonal term s), <t, s>, <t, t>
#pragma unroll
        for (int g = 0; g < 2 * G; g++) red[g] = 0;
        for (int c = threadIdx.x; c < N; c += BLOCK) {
            R sv_[G], off[G];
            each_band<G>(wm, [&](int g) { sv_[g] = V.r.ld(c, bg.so[g]); });
            group_offdiag<R, G>(bg, wm, V.sh, c, off);
            each_band<G>(wm, [&](int g) {
                const R tc = sv_[g] + bg.vt[g] * off[g];
                V.t.st(c, bg.so[g], tc);
                red[g] += (double)tc * (double)sv_[g];
                red[G + g] += (double)tc * (double)tc;
            });
        }
        block_sum<BLOCK, 2 * G>(red, lds);
        each_band<G>(wm, [&](int g) { s_omega[g] = (red[G + g] != 0) ? red[g] / red[G + g] : 0.0; });
        // (5) x += alpha p^ + omega s^, r = s - omega t, ||r||^2, <r_hat, r>
        {
            R al[G], om5[G];
            each_band<G>(wm, [&](int g) {
                al[g] = (R)s_alpha[g];
                om5[g] = (R)s_omega[g];
            });
#pragma unroll
            for (int g = 0; g < 2 * G; g++) red[g] = 0;
            for (int c = threadIdx.x; c < N; c += BLOCK) {
                R xv[G], phv[G], shv[G], rv[G], tv[G], rhv[G];
                each_band<G>(wm, [&](int g) {
                    const unsigned so = bg.so[g];
                    xv[g] = V.x.ld(c, so);
                    phv[g] = V.ph.ld(c, so);
                    shv[g] = V.sh.ld(c, so);
                    rv[g] = V.r.ld(c, so);
                    tv[g] = V.t.ld(c, so);
                    rhv[g] = V.rh.ld(c, so);
                });
                each_band<G>(wm, [&](int g) {
                    const unsigned so = bg.so[g];
                    V.x.st(c, so, xv[g] + (al[g] * phv[g] + om5[g] * shv[g]));
                    const R rn = rv[g] - om5[g] * tv[g];
                    V.r.st(c, so, rn);
                    red[g] += (double)rn * (double)rn;
                    red[G + g] += (double)rhv[g] * (double)rn;
                });
            }
        }
        block_sum<BLOCK, 2 * G>(red, lds);
        each_band<G>(wm, [&](int g) {
            const double nrm = sqrt(red[g]);
            s_nrm[g] = nrm;
            s_rnext[g] = red[G + g];
            if (nrm / s_nrm0[g] < tol) {
                its[g] = i;
                act &= ~(1u << g);
            }
        });
    }
    __syncthreads();
}

// Band-batched Jacobi inner solves (FP64): work-group item gid solves direction d = gid % D for
// bands b0 .. b0 + G - 1, b0 = (gid / D) G, on the right-hand sides in w.b.
template <int BLOCK, int G>
__global__ __launch_bounds__(BLOCK) void k_solve64_bands(Ops o, Work w, SolveCfg cfg) {
    __shared__ double lds[2 * G * (BLOCK / 64) + 7 * G];  // reductions, then the per-band scalars
    double* const st = lds + 2 * G * (BLOCK / 64);
    const int B = cfg.S / o.D;
    const int ngroups = o.D * ((B + G - 1) / G);
    for (int gid = blockIdx.x; gid < ngroups; gid += gridDim.x) {
        const BandGroup<double, G> bg = make_group<double, G>(o, B, gid % o.D, (gid / o.D) * G);
        const KrylovVecs<double> V{group_vec(bg, w.b),  group_vec(bg, w.x), group_vec(bg, w.r),
                                   group_vec(bg, w.rh), group_vec(bg, w.p), group_vec(bg, w.v),
                                   group_vec(bg, w.ph), group_vec(bg, w.sh), group_vec(bg, w.t)};
        int its[G];
        bicgstab_bands<double, BLOCK, G>(bg, bg.present, V, cfg.maxit, cfg.tol, cfg.reference_semantics != 0, lds,
                                         st, its);
        if (threadIdx.x == 0)
            each_band<G>(bg.present, [&](int g) {
                w.iters[bg.sys[g]] += (its[g] > cfg.maxit) ? cfg.maxit : its[g];
                w.fails[bg.sys[g]] += (its[g] > cfg.maxit) ? 1 : 0;
            });
        __syncthreads();
    }
}

// Band-batched mixed precision (Jacobi): k_solve_mixed's refinement for G bands of a direction,
// each band refined until its own FP64 residual criterion holds.
template <int BLOCK, int G>
__global__ __launch_bounds__(BLOCK) void k_solve_mixed_bands(Ops o, Work w, SolveCfg cfg) {
    // reductions, the inner solver's per-band scalars, then ||b|| and ||r64|| per band
    __shared__ double lds[2 * G * (BLOCK / 64) + 9 * G];
    double* const st = lds + 2 * G * (BLOCK / 64);
    double* const s_nb = st + 7 * G;
    double* const s_nr = st + 8 * G;
    const int B = cfg.S / o.D;
    const int ngroups = o.D * ((B + G - 1) / G);
    for (int gid = blockIdx.x; gid < ngroups; gid += gridDim.x) {
        const int d = gid % o.D, b0 = (gid / o.D) * G;
        const BandGroup<double, G> b64 = make_group<double, G>(o, B, d, b0);
        const BandGroup<float, G> b32 = make_group<float, G>(o, B, d, b0);
        const unsigned P = b64.present;
        const int N = o.N;
        const GVec<double> vb = group_vec(b64, w.b), vx = group_vec(b64, w.x), vr64 = group_vec(b64, w.r64);
        const KrylovVecs<float> V{group_vec(b32, w.f), group_vec(b32, w.y),  group_vec(b32, w.r),
                                  group_vec(b32, w.rh), group_vec(b32, w.p), group_vec(b32, w.v),
                                  group_vec(b32, w.ph), group_vec(b32, w.sh), group_vec(b32, w.t)};
        double red[G];
#pragma unroll
        for (int g = 0; g < G; g++) red[g] = 0;
        for (int c = threadIdx.x; c < N; c += BLOCK)
            each_band<G>(P, [&](int g) {
                const double bc = vb.ld(c, b64.so[g]);
                vr64.st(c, b64.so[g], bc);
                vx.st(c, b64.so[g], 0.0);
                red[g] += bc * bc;
            });
        block_sum<BLOCK, G>(red, lds);
        int total[G], fail[G];
#pragma unroll
        for (int g = 0; g < G; g++) {
            s_nb[g] = sqrt(red[g]);
            s_nr[g] = sqrt(red[g]);
            total[g] = fail[g] = 0;
        }
        for (int step = 0; step < cfg.refine_steps; step++) {
            unsigned rm = 0;
            each_band<G>(P, [&](int g) {
                if (s_nr[g] > cfg.tol * s_nb[g] && s_nr[g] > 1e-300) rm |= 1u << g;
            });
            if (!rm) break;
            double inv[G];
            each_band<G>(rm, [&](int g) { inv[g] = 1.0 / s_nr[g]; });
            for (int c = threadIdx.x; c < N; c += BLOCK)
                each_band<G>(rm, [&](int g) { V.rhs.st(c, b32.so[g], (float)(vr64.ld(c, b64.so[g]) * inv[g])); });
            __syncthreads();
            int its[G];
            bicgstab_bands<float, BLOCK, G>(b32, rm, V, cfg.maxit, 1e-6, false, lds, st, its);
            each_band<G>(rm, [&](int g) {
                total[g] += (its[g] > cfg.maxit) ? cfg.maxit : its[g];
                fail[g] += (its[g] > cfg.maxit) ? 1 : 0;
            });
            double nrv[G];
            each_band<G>(rm, [&](int g) { nrv[g] = s_nr[g]; });
            for (int c = threadIdx.x; c < N; c += BLOCK)
                each_band<G>(rm, [&](int g) {
                    vx.st(c, b64.so[g], vx.ld(c, b64.so[g]) + nrv[g] * (double)V.x.ld(c, b32.so[g]));
                });
            __syncthreads();
            // r64 = b - A x in FP64
#pragma unroll
            for (int g = 0; g < G; g++) red[g] = 0;
            for (int c = threadIdx.x; c < N; c += BLOCK) {
                const double dgc = buf_ld<double>(b64.dg, (unsigned)c * 8u, 0);
                const double vl = buf_ld<double>(b64.vol, (unsigned)c * 8u, 0);
                double xc[G], bc[G], off[G];
                each_band<G>(rm, [&](int g) {
                    xc[g] = vx.ld(c, b64.so[g]);
                    bc[g] = vb.ld(c, b64.so[g]);
                });
                group_offdiag<double, G>(b64, rm, vx, c, off);
                each_band<G>(rm, [&](int g) {
                    const double acc = group_diag(b64, g, dgc, vl) * xc[g];
                    const double rc = bc[g] - (acc + b64.vt[g] * off[g]);
                    vr64.st(c, b64.so[g], rc);
                    red[g] += rc * rc;
                });
            }
            block_sum<BLOCK, G>(red, lds);
            each_band<G>(rm, [&](int g) { s_nr[g] = sqrt(red[g]); });
        }
        if (threadIdx.x == 0)
            each_band<G>(P, [&](int g) {
                w.iters[b64.sys[g]] += total[g];
                w.fails[b64.sys[g]] += fail[g] + ((s_nr[g] > cfg.tol * s_nb[g] && s_nr[g] > 1e-300) ? 1 : 0);
            });
        __syncthreads();
    }
}

}  // namespace bte
