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

// Grow-only device buffer (released by its owner or on destruction).
struct DevBuf {
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void* p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

// Grow-only pinned host buffer.
struct HostBuf {
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() { release(); }
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
    const double* dUc = nullptr;  // or: device contact potentials (2 values)
    double t_odd = 0;             // gamma
    int sample[3] = {-1, -1, -1}; // integrand sample sites (KIND_NONEQ)
    int L = 0;                    // requested segment length (0: default)
    int chain0 = 0, nchain = 0;   // chains [chain0, chain0+nchain) of d_chains
    long long out_off = 0;        // offset of this tube's N outputs in d_out
    int ext = 0;                  // index of the tube for an active mask
};

// A reusable plan: tables and workspace sized once, enqueued many times.
class RgfPlan {
   public:
    ~RgfPlan();
    // kind: KIND_EQ / KIND_NONEQ / KIND_DOS.  d_chains: device ChainDesc array
    // (chains of tube t are [tubes[t].chain0, +nchain)); Sigma is written into it.
    int build(int kind, bool write_G, const std::vector<RgfTube>& tubes, ChainDesc* d_chains, int nchains,
              const DecParams& dec);
    // Enqueue: transfer matrices (with the lead self-energies), segment scan, sweeps and
    // the fixed-order row reduction.  d_out: per-tube outputs at out_off
    // (KIND_EQ/NONEQ); d_diagG: [chain][N]; d_chain_out: per-chain outputs
    // (integrand samples [slot][3], DOS/T pairs [slot][2]).
    // tube_active (device, indexed by RgfTube::ext, may be null): tubes with 0
    // are skipped and keep their previous outputs.  d_site_out (KIND_DOS,
    // optional, [chain][N]): each site's Gamma_L |G_n0|^2 + Gamma_R |G_n,N-1|^2.
    // finish = false leaves out the last step, the fixed-order sum of the
    // partial rows into d_out (KIND_EQ/NONEQ), for a caller that fuses it
    // into its next kernel (single_tube_rows).
    int enqueue(double* d_out, double2* d_diagG, double* d_chain_out, hipStream_t st,
                const int* tube_active = nullptr, double* d_site_out = nullptr, bool finish = true);
    // A one-tube plan's partial rows: row r of site n at part[r N + n], r <
    // nrows, summed in order r = 0, 1, ... from 0.0; false for other plans.
    bool single_tube_rows(const double** part, int* nrows) const
    {
        if (ntubes_ != 1 || kind_ == KIND_DOS) return false;
        *part = d_part_.as<const double>();
        *nrows = max_rows_;
        return true;
    }
    // Read and clear the device status word (synchronises the stream).
    int check_status(hipStream_t st, int* status);
    // The device status word itself (an int), for a copy enqueued by the caller.
    const int* status_ptr() const { return d_status_.as<int>(); }
    int num_items() const { return nitems_; }
    // hipEvent timing of the RGF kernels (transfer + scan + sweep)
    void enable_timing(bool on) { timing_ = on; }
    int kernel_time(double* total_ms, int* launches);

   private:
    int kind_ = 0;
    bool write_G_ = false;
    int nitems_ = 0, nchains_ = 0, ntubes_ = 0, nsblk_ = 0;
    bool any_fused_ = false, any_unfused_ = false;  // tubes with and without the fused transfer and scan
    bool any_fused2_ = false;                        // tubes whose equilibrium sweep is fused too
    std::vector<unsigned char> tube0_;               // kernel descriptor of the first tube (bytes)
    int max_N_ = 0, max_P_ = 0, max_groups_ = 0, max_rows_ = 0;
    ChainDesc* d_chains_ = nullptr;
    DecParams dec_;
    std::vector<RgfTube> tubes_;
    DevBuf d_tubes_, d_item_tube_, d_chain_tube_, d_sblk_tube_, d_sblk_chain_, d_item_list_;
    int fast_off_[NEGF_SUB_NS + 1] = {}, fast_n_[NEGF_SUB_NS + 1] = {};  // unrolled sweep items by sub-blocks
    int slow_off_ = 0, slow_n_ = 0;                                      // general sweep items
    DevBuf d_Tm_, d_St_, d_Dd_, d_part_, d_dos_part_, d_status_;
    DevBuf d_red_src_off_, d_red_nrows_, d_red_N_, d_red_dst_off_;  // fixed-order row reduction
    bool timing_ = false;
    std::vector<hipEvent_t> ev_start_, ev_stop_;
    size_t ev_used_ = 0;
};

// Host-side quadrature (Quadrature::Gauss_Legendre CodeUtil.cpp:640-684 and
// c_IntegrationPath::Define_GaussLegendrePoints IntegrationPath.cpp:41-147).
void gauss_legendre(int n, double* x, double* w);
void gl_path(double ar, double ai, double br, double bi, int n, int type, double* E, double* w, double* mul);

// Segmentation of a chain of N sites into P segments of L sites (L a
// multiple of NEGF_SUB_K, at most NEGF_SEG_MAX); 0 if N is too long.
int choose_segments(int N, int L_req, int* P, int* L);
// Segment length for a lone tube when none is asked for: a lone tube has one
// chain per energy, a few chain groups, so its launch is short of wavefronts;
// where 128-site segments would give fewer segments than the fused transfer
// and walk launch takes, the shortest multiple of NEGF_SUB_K that keeps the
// count within that launch's range.  0 (the default) otherwise.
int lone_tube_segment_length(int N);

// Batched solve internals shared with the batched SCF: the host part of a
// solve (rebuilding plans when paths changed) and the bare launch sequence
// on a given stream (capturable).
int batch_prepare(negf_batch* b, int iter, bool* rebuilt);
int batch_enqueue(negf_batch* b, double* dQ, hipStream_t st, const int* tube_active = nullptr);
// Build counter of the batch's plans: a graph captured over batch_enqueue
// is valid only while this value is unchanged.
long long batch_generation(const negf_batch* b);
// Fold the plans' device status words into the batch's sticky root flag
// (reported and cleared by negf_batch_status); synchronises.
int batch_poll_status(negf_batch* b);
// The device status words of the plans batch_enqueue launches (neq: nullptr
// when the batch has no non-equilibrium plan); valid until the next rebuild.
void batch_status_words(const negf_batch* b, const int** eq, const int** neq);
int max_sites();

}  // namespace negf
