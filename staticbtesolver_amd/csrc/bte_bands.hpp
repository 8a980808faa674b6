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

// Off-diagonal sums of the G bands' products at the K cells c0 + k BLOCK: off[k][g] = sum_f a_f
// x_g[j_f], each row loaded once.  Slots beyond the width and missing neighbours (column -1)
// contribute nothing; the gather of a missing neighbour reads the cell itself, so no load is
// conditional.
template <typename R, int G, int BLOCK, int K>
__device__ __forceinline__ void group_offdiag(const BandGroup<R, G>& bg, unsigned m, const GVec<R>& x, int c0,
                                              R (&off)[K][G]) {
    const int W = bg.W;
    int j[K][3];
    R a[K][3];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const unsigned c = (unsigned)(c0 + k * BLOCK);
#pragma unroll
        for (int f = 0; f < 3; f++) {
            const unsigned fo = (unsigned)f * (unsigned)bg.N;
            j[k][f] = (f < W) ? buf_ld<int>(bg.col, c * 4u, fo * 4u) : -1;
            a[k][f] = (f < W) ? buf_ld<R>(bg.val, c * (unsigned)sizeof(R), fo * (unsigned)sizeof(R)) : (R)0;
        }
    }
    R gx[K][G][3];
#pragma unroll
    for (int k = 0; k < K; k++)
        each_band<G>(m, [&](int g) {
#pragma unroll
            for (int f = 0; f < 3; f++) gx[k][g][f] = x.ld(j[k][f] >= 0 ? j[k][f] : c0 + k * BLOCK, bg.so[g]);
        });
#pragma unroll
    for (int k = 0; k < K; k++)
        each_band<G>(m, [&](int g) {
            R o = 0;
#pragma unroll
            for (int f = 0; f < 3; f++)
                if (j[k][f] >= 0) o += a[k][f] * gx[k][g][f];
            off[k][g] = o;
        });
}

template <typename R, int G>
__device__ __forceinline__ R group_diag(const BandGroup<R, G>& bg, int g, R dgc, double vl) {
    return bg.vt[g] * dgc + bg.wv * (R)vl;
}

// Cells per thread and step of the band-batched passes: about 8 (FP64: 4) values per streamed
// array in flight per lane, over cells and bands together.
template <typename R, int G>
constexpr int band_unroll() {
    return ((sizeof(R) == 4 ? 8 : 4) / G) > 0 ? ((sizeof(R) == 4 ? 8 : 4) / G) : 1;
}

template <unsigned M>
struct MaskC {
    static constexpr unsigned value = M;
};

// for_cells over the bands in m: the loop is instantiated for the full mask (and, for G = 2, for
// each single band) with the mask a compile-time constant, so the per-band tests fold away and the
// loads of all bands and cells of a step issue back to back; other masks run a generic copy whose
// per-band branches split the loads.  body(kc, c0, mc): mc::value is the mask, 0 in the generic copy.
template <int BLOCK, int U, int G, class F>
__device__ __forceinline__ void band_cells(int N, unsigned m, F&& body) {
    constexpr unsigned full = (1u << G) - 1;
    auto go = [&](auto mc) { for_cells<BLOCK, U>(N, [&](auto kc, int c0) { body(kc, c0, mc); }); };
    if (m == full) {
        go(MaskC<full>{});
        return;
    }
    if constexpr (G == 2) {
        if (m == 1) {
            go(MaskC<1>{});
            return;
        }
        if (m == 2) {
            go(MaskC<2>{});
            return;
        }
    }
    go(MaskC<0>{});
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
    constexpr int U = band_unroll<R, G>();
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
    band_cells<BLOCK, U, G>(N, act, [&](auto kc, int c0, auto mc) {
        const unsigned msk = decltype(mc)::value ? decltype(mc)::value : act;
        constexpr int K = decltype(kc)::value;
        R bc[K][G];
#pragma unroll
        for (int k = 0; k < K; k++) each_band<G>(msk, [&](int g) { bc[k][g] = V.rhs.ld(c0 + k * BLOCK, bg.so[g]); });
#pragma unroll
        for (int k = 0; k < K; k++)
            each_band<G>(msk, [&](int g) {
                const int c = c0 + k * BLOCK;
                const unsigned so = bg.so[g];
                V.x.st(c, so, (R)0);
                V.r.st(c, so, bc[k][g]);
                V.rh.st(c, so, bc[k][g]);
                V.p.st(c, so, refsem ? bc[k][g] : (R)0);
                V.v.st(c, so, (R)0);
                red[g] += (double)bc[k][g] * (double)bc[k][g];
            });
    });
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
        band_cells<BLOCK, U, G>(N, wm, [&](auto kc, int c0, auto mc) {
            const unsigned msk = decltype(mc)::value ? decltype(mc)::value : wm;
            constexpr int K = decltype(kc)::value;
            R dgc[K], pv[K][G], vv[K][G], rv[K][G];
            double vl[K];
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int c = c0 + k * BLOCK;
                dgc[k] = buf_ld<R>(bg.dg, (unsigned)c * sizeof(R), 0);
                vl[k] = buf_ld<double>(bg.vol, (unsigned)c * 8u, 0);
                each_band<G>(msk, [&](int g) {
                    pv[k][g] = V.p.ld(c, bg.so[g]);
                    vv[k][g] = V.v.ld(c, bg.so[g]);
                    rv[k][g] = V.r.ld(c, bg.so[g]);
                });
            }
#pragma unroll
            for (int k = 0; k < K; k++)
                each_band<G>(msk, [&](int g) {
                    const int c = c0 + k * BLOCK;
                    const R pn = (pv[k][g] - om[g] * vv[k][g]) * be[g] + rv[k][g];
                    V.p.st(c, bg.so[g], pn);
                    V.ph.st(c, bg.so[g], pn / group_diag(bg, g, dgc[k], vl[k]));
                });
        });
        __syncthreads();
        // (2) v = A p^ (diagonal term p), <r_hat, v>
#pragma unroll
        for (int g = 0; g < G; g++) red[g] = 0;
        band_cells<BLOCK, U, G>(N, wm, [&](auto kc, int c0, auto mc) {
            const unsigned msk = decltype(mc)::value ? decltype(mc)::value : wm;
            constexpr int K = decltype(kc)::value;
            R dz[K][G], rhv[K][G], off[K][G];
#pragma unroll
            for (int k = 0; k < K; k++)
                each_band<G>(msk, [&](int g) {
                    dz[k][g] = V.p.ld(c0 + k * BLOCK, bg.so[g]);
                    rhv[k][g] = V.rh.ld(c0 + k * BLOCK, bg.so[g]);
                });
            group_offdiag<R, G, BLOCK, K>(bg, msk, V.ph, c0, off);
#pragma unroll
            for (int k = 0; k < K; k++)
                each_band<G>(msk, [&](int g) {
                    const R vc = dz[k][g] + bg.vt[g] * off[k][g];
                    V.v.st(c0 + k * BLOCK, bg.so[g], vc);
                    red[g] += (double)rhv[k][g] * (double)vc;
                });
        });
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
            band_cells<BLOCK, U, G>(N, wm, [&](auto kc, int c0, auto mc) {
                const unsigned msk = decltype(mc)::value ? decltype(mc)::value : wm;
                constexpr int K = decltype(kc)::value;
                R dgc[K], rv[K][G], vv[K][G];
                double vl[K];
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const int c = c0 + k * BLOCK;
                    dgc[k] = buf_ld<R>(bg.dg, (unsigned)c * sizeof(R), 0);
                    vl[k] = buf_ld<double>(bg.vol, (unsigned)c * 8u, 0);
                    each_band<G>(msk, [&](int g) {
                        rv[k][g] = V.r.ld(c, bg.so[g]);
                        vv[k][g] = V.v.ld(c, bg.so[g]);
                    });
                }
#pragma unroll
                for (int k = 0; k < K; k++)
                    each_band<G>(msk, [&](int g) {
                        const int c = c0 + k * BLOCK;
                        const R s = rv[k][g] - al[g] * vv[k][g];
                        V.r.st(c, bg.so[g], s);
                        V.sh.st(c, bg.so[g], s / group_diag(bg, g, dgc[k], vl[k]));
                        red[g] += (double)s * (double)s;
                    });
            });
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
        band_cells<BLOCK, U, G>(N, wm, [&](auto kc, int c0, auto mc) {
            const unsigned msk = decltype(mc)::value ? decltype(mc)::value : wm;
            constexpr int K = decltype(kc)::value;
            R sv_[K][G], off[K][G];
#pragma unroll
            for (int k = 0; k < K; k++) each_band<G>(msk, [&](int g) { sv_[k][g] = V.r.ld(c0 + k * BLOCK, bg.so[g]); });
            group_offdiag<R, G, BLOCK, K>(bg, msk, V.sh, c0, off);
#pragma unroll
            for (int k = 0; k < K; k++)
                each_band<G>(msk, [&](int g) {
                    const R tc = sv_[k][g] + bg.vt[g] * off[k][g];
                    V.t.st(c0 + k * BLOCK, bg.so[g], tc);
                    red[g] += (double)tc * (double)sv_[k][g];
                    red[G + g] += (double)tc * (double)tc;
                });
        });
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
            band_cells<BLOCK, U, G>(N, wm, [&](auto kc, int c0, auto mc) {
                const unsigned msk = decltype(mc)::value ? decltype(mc)::value : wm;
                constexpr int K = decltype(kc)::value;
                R xv[K][G], phv[K][G], shv[K][G], rv[K][G], tv[K][G], rhv[K][G];
#pragma unroll
                for (int k = 0; k < K; k++)
                    each_band<G>(msk, [&](int g) {
                        const int c = c0 + k * BLOCK;
                        const unsigned so = bg.so[g];
                        xv[k][g] = V.x.ld(c, so);
                        phv[k][g] = V.ph.ld(c, so);
                        shv[k][g] = V.sh.ld(c, so);
                        rv[k][g] = V.r.ld(c, so);
                        tv[k][g] = V.t.ld(c, so);
                        rhv[k][g] = V.rh.ld(c, so);
                    });
#pragma unroll
                for (int k = 0; k < K; k++)
                    each_band<G>(msk, [&](int g) {
                        const int c = c0 + k * BLOCK;
                        const unsigned so = bg.so[g];
                        V.x.st(c, so, xv[k][g] + (al[g] * phv[k][g] + om5[g] * shv[k][g]));
                        const R rn = rv[k][g] - om5[g] * tv[k][g];
                        V.r.st(c, so, rn);
                        red[g] += (double)rn * (double)rn;
                        red[G + g] += (double)rhv[k][g] * (double)rn;
                    });
            });
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

}  // namespace bte
