// bte_internal.hpp -- host-side types shared by the solver (bte_lib.hip) and the graph-network
// (bte_gnn.hip) translation units of libbte_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

namespace bte {

// Set (per thread) while an entry point that must not allocate device memory runs: a DBuf that
// would have to grow then fails with hipErrorOutOfMemory instead of calling hipMalloc, and the
// flag records that it happened.
struct AllocGuard {
    static bool& on() {
        static thread_local bool f = false;
        return f;
    }
    static bool& tripped() {
        static thread_local bool t = false;
        return t;
    }
};

struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t ensure(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        if (AllocGuard::on()) {
            AllocGuard::tripped() = true;
            return hipErrorOutOfMemory;
        }
        release();
        hipError_t e = hipMalloc(&p, std::max<size_t>(n, 16));
        if (e == hipSuccess) bytes = n;
        return e;
    }
    template <class T>
    T* as() const {
        return (T*)p;
    }
};

constexpr int kGnnIn = 16;   // input features per node (zero padded)
constexpr int kGnnHid = 64;  // hidden width

// Mesh graph network state on one device: parameters (FP32 master copy; GEMM operands are rounded
// to BF16 at use), Adam moments, graph, activations kept for the backward pass.
struct GnnNet {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int L = 4;       // message-passing layers
    int P = 0;       // number of parameters
    int N = 0, F = 0;
    int cap = 0;                        // node capacity of the activation buffers
    unsigned long long graph_tag = 0;   // geometry tag of the solver mesh the graph holds (0: none)
    DBuf params, grads, adam_m, adam_v;
    int adam_t = 0;
    DBuf nbr, deg;                      // graph (face table, interior degree)
    DBuf X, y, tgt, wgt, dy;            // inputs / outputs of the last pass
    DBuf Z0;                            // encoder pre-activation
    std::vector<DBuf> H, M, Z;          // H[0..L], M[0..L-1], Z[0..L-1]
    DBuf dH, dH2, dZ, G1, G2, part, red;
    size_t part_elems = 0;
    // inference: the message-passing weights as BF16 MFMA operand fragments (k_gnn_wfrag), built
    // from the parameters of version wfrag_ver; params_ver counts parameter changes
    DBuf wfrag;
    unsigned long long params_ver = 1, wfrag_ver = 0;
};

// Offsets of the parameter blocks in the flat vector (layout shared with oracle/gnn_ref.py).
__host__ __device__ inline int gnn_off_layer(int l) {
    return kGnnIn * kGnnHid + kGnnHid + l * (2 * kGnnHid * kGnnHid + kGnnHid);
}
inline int gnn_off_out(int L) { return gnn_off_layer(L); }
inline int gnn_num_params(int L) { return gnn_off_out(L) + kGnnHid + 1; }

// Forward pass on device data already in net->X (N rows); writes net->y, and keeps the activations
// the backward pass reads.  Launches on `stream` (the network's own stream when null).  Returns 0
// or a negative bte_status.
int gnn_forward_device(GnnNet* net, hipStream_t stream = nullptr);
// The same output y without the backward pass's activations: one fused launch per layer (mean
// aggregation, both products, residual; the decoder in the last), on BF16 operand fragments of
// the weights.  Used by the solver's learned correction and bte_gnn_forward.
int gnn_infer_device(GnnNet* net, hipStream_t stream = nullptr);
// Ensures activation buffers for n nodes.
int gnn_ensure(GnnNet* net, int n);

}  // namespace bte

// The opaque handle of the C ABI is the network state itself.
struct bte_gnn : bte::GnnNet {};
