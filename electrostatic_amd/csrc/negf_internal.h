// negf_internal.h -- host-side internals shared by the library translation
// units: error reporting, device workspace and the batched RGF launcher.
#pragma once
#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "../../include/negf_amd.h"
#include "negf_types.h"

namespace negf {

void set_error(const char* fmt, ...);
void clear_error();

#define NEGF_HIP_TRY(expr)                                                                  \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess) {                                                             \
            ::negf::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                              __LINE__);                                                    \
            return NEGF_EDEVICE;                                                            \
        }                                                                                   \
    } while (0)

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Grow-only pinned host buffer.
struct HostBuf {
    void* p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

struct DecParams {
    int use = 0, layers = 10, max_iter = 30, rel_error = 0;
};

// One tube of a batched RGF launch.
struct RgfTube {
    int N = 0;
    const double* dU = nullptr;   // device potential, N values
    double Uc0 = 0, Uc1 = 0;      // contact potentials
    double t_odd = 0;             // gamma
    int sample[3] = {-1, -1, -1}; // integrand sample sites (KIND_NONEQ)
    int chain0 = 0, nchain = 0;   // chains [chain0, chain0+nchain) of d_chains
    long long out_off = 0;        // offset of this tube's N outputs in d_out
};

// A reusable plan: block tables and workspace sized once, enqueued many times.
class RgfPlan {
   public:
    ~RgfPlan();
    // kind: KIND_EQ / KIND_NONEQ / KIND_DOS.  d_chains: device ChainDesc array.
    int build(int kind, bool write_G, const std::vector<RgfTube>& tubes, const ChainDesc* d_chains,
              const DecParams& dec);
    // Enqueue: potential transpose, RGF kernel(s), fixed-order row reduction.
    // d_out: per-tube outputs at out_off (KIND_EQ/NONEQ); d_diagG: [chain][N];
    // d_chain_out: per-chain outputs (integrand samples / DOS pairs).
    int enqueue(double* d_out, double2* d_diagG, double* d_chain_out, hipStream_t st);
    // Read and clear the device status word (synchronises the stream).
    int check_status(hipStream_t st, int* status);
    int num_blocks() const { return nblocks_; }
    int variant() const { return variant_; }

   private:
    int kind_ = 0;
    bool write_G_ = false;
    int variant_ = 0;
    int nblocks_ = 0;
    int ntubes_ = 0;
    int lds_bytes_ = 0;
    bool any_global_atomic_ = false;
    const ChainDesc* d_chains_ = nullptr;
    DecParams dec_;
    std::vector<RgfTube> tubes_;
    std::vector<int> Lf_;
    std::vector<long long> u_off_;
    long long part_elems_ = 0;
    int reduce_chunk_ = 16;
    int max_rows_ = 0, max_N_ = 0, max_chunks_ = 0;
    long long tmp_elems_ = 0;
    DevBuf d_tubes_, d_bchain0_, d_bnchain_, d_btube_, d_bpart_;
    DevBuf d_Ut_, d_part_, d_tmp_, d_status_;
    DevBuf d_red_src_off_, d_red_nrows_, d_red_N_, d_red_dst_off_;   // pass 1
    DevBuf d_red2_src_off_, d_red2_nrows_, d_red2_dst_off_;          // pass 2
};

// Choose the kernel variant for a chain of N sites: returns the full segment
// length Lf (multiple of the variant's sub-block), or 0 if N is too long.
int choose_variant(int N, int* variant, int* Lf);
int max_sites();

}  // namespace negf
