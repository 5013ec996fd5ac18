// negf_engine.hip -- host launcher for the segmented RGF kernels and the
// Level-1 C ABI (batched RGF over an energy list for one chain).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "negf_internal.h"
#include "rgf_wave.hip"

namespace negf {

// ---------------------------------------------------------------- errors
static thread_local std::string g_err;

void set_error(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}
void clear_error() { g_err.clear(); }

// ---------------------------------------------------------------- buffers
int DevBuf::ensure(size_t bytes)
{
    if (bytes <= cap) return NEGF_OK;
    release();
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        set_error("hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
        return NEGF_EALLOC;
    }
    cap = want;
    return NEGF_OK;
}
void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}
int HostBuf::ensure(size_t bytes)
{
    if (bytes <= cap) return NEGF_OK;
    release();
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        p = nullptr;
        set_error("hipHostMalloc(%zu) failed: %s", want, hipGetErrorString(e));
        return NEGF_EALLOC;
    }
    cap = want;
    return NEGF_OK;
}
void HostBuf::release()
{
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
}

// ---------------------------------------------------------------- segmentation
// Segments of L sites (a multiple of the sub-block K, at most NEGF_SEG_MAX);
// all but the last are full.  L_req = 0 picks the default.
int choose_segments(int N, int L_req, int* P, int* L)
{
    if (N < 1) return 0;
    int l = L_req > 0 ? L_req : NEGF_SEG_DEFAULT;
    l = (l + NEGF_SUB_K - 1) / NEGF_SUB_K * NEGF_SUB_K;
    l = std::min(std::max(l, NEGF_SUB_K), NEGF_SEG_MAX);
    l = std::min(l, (N + NEGF_SUB_K - 1) / NEGF_SUB_K * NEGF_SUB_K);
    const long long p = (N + l - 1) / l;
    if (p > NEGF_MAX_SEGMENTS) return 0;
    *P = (int)p;
    *L = l;
    return 1;
}
int max_sites() { return NEGF_MAX_SITES; }

template <int KIND, bool WG>
static void launch_sweep(int nblk, const WLaunch& lp, hipStream_t st)
{
    hipLaunchKernelGGL((sweep_kernel<NEGF_SUB_K, NEGF_SUB_NS, KIND, WG>), dim3(nblk), dim3(256), 0, st, lp);
}

static void launch_sweep_neq_full(int ns, int nblk, const WLaunch& lp, hipStream_t st)
{
    constexpr int K = NEGF_SUB_K;
    switch (ns) {
        case 1: hipLaunchKernelGGL((sweep_neq_full_kernel<K, 1>), dim3(nblk), dim3(256), 0, st, lp); break;
        case 2: hipLaunchKernelGGL((sweep_neq_full_kernel<K, 2>), dim3(nblk), dim3(256), 0, st, lp); break;
        case 3: hipLaunchKernelGGL((sweep_neq_full_kernel<K, 3>), dim3(nblk), dim3(256), 0, st, lp); break;
        case 4: hipLaunchKernelGGL((sweep_neq_full_kernel<K, 4>), dim3(nblk), dim3(256), 0, st, lp); break;
        case 5: hipLaunchKernelGGL((sweep_neq_full_kernel<K, 5>), dim3(nblk), dim3(256), 0, st, lp); break;
        case 6: hipLaunchKernelGGL((sweep_neq_full_kernel<K, 6>), dim3(nblk), dim3(256), 0, st, lp); break;
        case 7: hipLaunchKernelGGL((sweep_neq_full_kernel<K, 7>), dim3(nblk), dim3(256), 0, st, lp); break;
        default: hipLaunchKernelGGL((sweep_neq_full_kernel<K, 8>), dim3(nblk), dim3(256), 0, st, lp); break;
    }
}

static void launch_sweep_eq_full(int ns, int nblk, const WLaunch& lp, hipStream_t st)
{
    constexpr int K = NEGF_SUB_K;
    switch (ns) {
        case 1: hipLaunchKernelGGL((sweep_eq_full_kernel<K, 1>), dim3(nblk), dim3(256), 0, st, lp); break;
        case 2: hipLaunchKernelGGL((sweep_eq_full_kernel<K, 2>), dim3(nblk), dim3(256), 0, st, lp); break;
        case 3: hipLaunchKernelGGL((sweep_eq_full_kernel<K, 3>), dim3(nblk), dim3(256), 0, st, lp); break;
        case 4: hipLaunchKernelGGL((sweep_eq_full_kernel<K, 4>), dim3(nblk), dim3(256), 0, st, lp); break;
        case 5: hipLaunchKernelGGL((sweep_eq_full_kernel<K, 5>), dim3(nblk), dim3(256), 0, st, lp); break;
        case 6: hipLaunchKernelGGL((sweep_eq_full_kernel<K, 6>), dim3(nblk), dim3(256), 0, st, lp); break;
        case 7: hipLaunchKernelGGL((sweep_eq_full_kernel<K, 7>), dim3(nblk), dim3(256), 0, st, lp); break;
        default: hipLaunchKernelGGL((sweep_eq_full_kernel<K, 8>), dim3(nblk), dim3(256), 0, st, lp); break;
    }
}

// ---------------------------------------------------------------- plan
RgfPlan::~RgfPlan()
{
    DevBuf* bufs[] = {&d_item_list_, &d_tubes_,    &d_item_tube_, &d_chain_tube_,  &d_sblk_tube_,
                      &d_sblk_chain_, &d_Tm_,       &d_St_,        &d_Dd_,          &d_part_,
                      &d_dos_part_,   &d_status_,   &d_red_src_off_, &d_red_nrows_, &d_red_N_,
                      &d_red_dst_off_};
    for (DevBuf* b : bufs) b->release();
    for (hipEvent_t e : ev_start_) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_stop_) (void)hipEventDestroy(e);
}

int RgfPlan::kernel_time(double* total_ms, int* launches)
{
    double tot = 0.0;
    for (size_t i = 0; i < ev_used_; ++i) {
        NEGF_HIP_TRY(hipEventSynchronize(ev_stop_[i]));
        float ms = 0.f;
        NEGF_HIP_TRY(hipEventElapsedTime(&ms, ev_start_[i], ev_stop_[i]));
        tot += ms;
    }
    *total_ms = tot;
    *launches = (int)ev_used_;
    ev_used_ = 0;
    return NEGF_OK;
}

// Tubes of FUSED_MIN_SEGS..FUSED_SEGS segments take the fused transfer and
// scan launch (NEGF_RGF_FUSED=0 in the environment turns it off, which leaves
// the results the same up to rounding).
static int fused_for(int P)
{
    const char* e = std::getenv("NEGF_RGF_FUSED");
    if (e && e[0] == '0') return 0;
    return P >= FUSED_MIN_SEGS && P <= FUSED_SEGS ? 1 : 0;
}

// Kernel-side descriptor of one tube whose items start at item0 and whose
// partial rows start at part_off.
int lone_tube_segment_length(int N)
{
    if ((N + NEGF_SEG_DEFAULT - 1) / NEGF_SEG_DEFAULT >= FUSED_MIN_SEGS) return 0;
    const int L = (N + FUSED_SEGS - 1) / FUSED_SEGS;
    const int Lk = (L + NEGF_SUB_K - 1) / NEGF_SUB_K * NEGF_SUB_K;
    return (N + Lk - 1) / Lk >= FUSED_MIN_SEGS ? Lk : 0;
}

static WTube make_wtube(const RgfTube& tb, int P, int L, int item0, long long part_off)
{
    WTube w;
    std::memset(&w, 0, sizeof(w));
    w.N = tb.N;
    w.L = L;
    w.P = P;
    w.nsub = L / NEGF_SUB_K;
    w.t_odd = tb.t_odd;
    w.Uc0 = tb.Uc0;
    w.Uc1 = tb.Uc1;
    w.U = tb.dU;
    w.Ucp = tb.dUc;
    w.chain0 = tb.chain0;
    w.nchain = tb.nchain;
    w.item0 = item0;
    w.part_off = part_off;
    for (int q = 0; q < 3; ++q) w.sample[q] = tb.sample[q];
    w.ext = tb.ext;
    w.fused = fused_for(P);
    return w;
}

template <class T>
static int upload(DevBuf& b, const std::vector<T>& v)
{
    int rc = b.ensure(std::max<size_t>(1, v.size()) * sizeof(T));
    if (rc) return rc;
    if (!v.empty()) NEGF_HIP_TRY(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return NEGF_OK;
}

int RgfPlan::build(int kind, bool write_G, const std::vector<RgfTube>& tubes, ChainDesc* d_chains, int nchains,
                   const DecParams& dec)
{
    if (tubes.empty() || !d_chains || nchains <= 0) {
        set_error("RgfPlan::build: no tubes or chains");
        return NEGF_EARG;
    }
    if (dec.use && (dec.layers < 2 || dec.layers > NEGF_MAX_DECIMATION_LAYERS)) {
        set_error("decimation_layers must be in [2, %d]", NEGF_MAX_DECIMATION_LAYERS);
        return NEGF_EARG;
    }
    kind_ = kind;
    write_G_ = write_G;
    d_chains_ = d_chains;
    nchains_ = nchains;
    dec_ = dec;
    tubes_ = tubes;
    ntubes_ = (int)tubes.size();
    std::vector<WTube> wt(ntubes_);
    std::vector<int> item_tube, chain_tube(nchains, 0), sblk_tube, sblk_chain;
    std::vector<int> fast_items[NEGF_SUB_NS + 1], slow_items;
    std::vector<long long> rsrc, rdst;
    std::vector<int> rnrows, rN;
    long long poff = 0;
    int items = 0;
    max_N_ = 0;
    max_rows_ = 0;
    max_P_ = 0;
    max_groups_ = 0;
    any_fused_ = any_fused2_ = any_unfused_ = false;
    for (int t = 0; t < ntubes_; ++t) {
        const RgfTube& tb = tubes[t];
        if (tb.N < 2 || !tb.dU || tb.nchain <= 0 || tb.t_odd <= 0 || tb.chain0 < 0 ||
            tb.chain0 + tb.nchain > nchains) {
            set_error("RgfPlan::build: invalid tube %d (N=%d, chains %d+%d of %d)", t, tb.N, tb.chain0, tb.nchain,
                      nchains);
            return NEGF_EARG;
        }
        int P, L;
        if (!choose_segments(tb.N, tb.L, &P, &L)) {
            set_error("chain of %d sites exceeds the supported maximum %d", tb.N, max_sites());
            return NEGF_ENOTSUP;
        }
        WTube& w = wt[t];
        w = make_wtube(tb, P, L, items, poff);
        // equilibrium charge on full 128-site segments: the whole launch fused
        if (w.fused && kind == KIND_EQ && !write_G && L == NEGF_SEG_MAX && tb.N % L == 0) w.fused = 2;
        (w.fused == 2 ? any_fused2_ : w.fused ? any_fused_ : any_unfused_) = true;
        const int G = (tb.nchain + 63) / 64;
        const int rows = w.fused == 2 ? 2 * G : G;  // partial rows: one per 32 chains when fused
        for (int i = 0; i < G * P; ++i) {
            item_tube.push_back(t);
            if (w.fused == 2) continue;  // no separate sweep
            // full segments of the equilibrium charge go to the unrolled sweep
            const int sgi = i % P;
            const bool full = sgi * L + L <= tb.N;
            const bool samples = tb.sample[0] >= 0 || tb.sample[1] >= 0 || tb.sample[2] >= 0;
            if (full && ((kind == KIND_EQ && !write_G) || (kind == KIND_NONEQ && !samples)))
                fast_items[L / NEGF_SUB_K].push_back(items + i);
            else
                slow_items.push_back(items + i);
        }
        items += G * P;
        for (int c = 0; c < tb.nchain; ++c) chain_tube[tb.chain0 + c] = t;
        w.group0 = (int)sblk_tube.size();
        for (int g = 0; g < G; ++g) {  // one scan wavefront per chain group
            sblk_tube.push_back(t);
            sblk_chain.push_back(g);
        }
        rsrc.push_back(poff);
        rnrows.push_back(rows);
        rN.push_back(tb.N);
        poff += (long long)rows * tb.N;
        rdst.push_back(tb.out_off);
        max_N_ = std::max(max_N_, tb.N);
        max_rows_ = std::max(max_rows_, rows);
        max_P_ = std::max(max_P_, P);
        max_groups_ = std::max(max_groups_, G);
    }
    nitems_ = items;
    nsblk_ = (int)sblk_tube.size();
    int rc;
    if ((rc = d_Tm_.ensure((size_t)items * 9 * 64 * sizeof(double)))) return rc;
    if ((rc = d_St_.ensure((size_t)items * 10 * 64 * sizeof(double)))) return rc;
    if ((rc = d_Dd_.ensure((size_t)nchains * 3 * sizeof(double)))) return rc;
    if ((rc = d_part_.ensure(std::max<long long>(1, poff) * sizeof(double)))) return rc;
    if (kind == KIND_DOS && (rc = d_dos_part_.ensure((size_t)items * 64 * sizeof(double)))) return rc;
    if ((rc = upload(d_tubes_, wt))) return rc;
    tube0_.assign(reinterpret_cast<const unsigned char*>(wt.data()),
                  reinterpret_cast<const unsigned char*>(wt.data()) + sizeof(WTube));
    if ((rc = upload(d_item_tube_, item_tube))) return rc;
    {
        std::vector<int> all;
        for (int ns = 1; ns <= NEGF_SUB_NS; ++ns) {
            fast_off_[ns] = (int)all.size();
            fast_n_[ns] = (int)fast_items[ns].size();
            all.insert(all.end(), fast_items[ns].begin(), fast_items[ns].end());
        }
        slow_off_ = (int)all.size();
        slow_n_ = (int)slow_items.size();
        all.insert(all.end(), slow_items.begin(), slow_items.end());
        if ((rc = upload(d_item_list_, all))) return rc;
    }
    if ((rc = upload(d_chain_tube_, chain_tube))) return rc;
    if ((rc = upload(d_sblk_tube_, sblk_tube))) return rc;
    if ((rc = upload(d_sblk_chain_, sblk_chain))) return rc;
    if ((rc = upload(d_red_src_off_, rsrc))) return rc;
    if ((rc = upload(d_red_nrows_, rnrows))) return rc;
    if ((rc = upload(d_red_N_, rN))) return rc;
    if ((rc = upload(d_red_dst_off_, rdst))) return rc;
    if ((rc = d_status_.ensure(sizeof(int)))) return rc;
    NEGF_HIP_TRY(hipMemset(d_status_.p, 0, sizeof(int)));
    return NEGF_OK;
}

// Everything one launch sequence reads: device tables and workspace, and the
// host counts that size the launches.  A plan fills it from its own buffers;
// the no-allocation entry point from the fixed scratch array.
struct RgfLayout {
    int kind = 0;
    bool write_G = false;
    int nitems = 0, nchains = 0, ntubes = 0, nsblk = 0, max_N = 0;
    int max_rows = 1 << 30;  // most partial rows of a tube (the reduction's form)
    ChainDesc* chains = nullptr;
    const WTube* tubes = nullptr;
    const int* item_tube = nullptr;
    const int* chain_tube = nullptr;
    const int* sblk_tube = nullptr;
    const int* sblk_chain = nullptr;
    const int* item_list = nullptr;  // fast items by sub-block count, then the general-sweep items
    int fast_off[NEGF_SUB_NS + 1] = {}, fast_n[NEGF_SUB_NS + 1] = {};
    int slow_off = 0, slow_n = 0;
    double *Tm = nullptr, *St = nullptr, *Dd = nullptr, *part = nullptr, *dos_part = nullptr;
    int* status = nullptr;
    const long long* red_src_off = nullptr;  // fixed-order row reduction tables
    const int* red_nrows = nullptr;
    const int* red_N = nullptr;
    const long long* red_dst_off = nullptr;
    DecParams dec;
    bool any_fused = false, any_unfused = true;  // tubes with and without the fused transfer and scan
    bool any_fused2 = false;                     // tubes whose equilibrium sweep is fused too
    const WTube* tube1 = nullptr;                // host copy of the descriptor when ntubes == 1
};

static WLaunch launch_params(const RgfLayout& y, double2* d_diagG, double* d_chain_out, const int* tube_active,
                             double* d_site_out = nullptr)
{
    WLaunch lp;
    std::memset(&lp, 0, sizeof(lp));
    lp.chains = y.chains;
    lp.tubes = y.tubes;
    lp.item_tube = y.item_tube;
    lp.nitems = y.nitems;
    lp.Tm = y.Tm;
    lp.St = y.St;
    lp.Dd = y.Dd;
    lp.part = y.part;
    lp.diagG = d_diagG;
    lp.chain_out = d_chain_out;
    lp.dos_part = y.dos_part;
    lp.sblk_tube = y.sblk_tube;
    lp.single = y.ntubes == 1 && y.tube1 ? 1 : 0;
    if (lp.single) lp.tube1 = *y.tube1;
    lp.sblk_chain = y.sblk_chain;
    lp.status = y.status;
    lp.kind = y.kind;
    lp.dec_use = y.dec.use;
    lp.dec_layers = y.dec.layers;
    lp.dec_max_iter = y.dec.max_iter;
    lp.dec_rel_error = y.dec.rel_error;
    lp.tube_active = tube_active;
    lp.site_out = d_site_out;
    return lp;
}

// true when the layout is one launch of rgf_eq_fused_kernel and nothing else
static bool single_fused_launch(const RgfLayout& y)
{
    if (!y.any_fused2 || y.any_fused || y.any_unfused || y.slow_n > 0) return false;
    for (int ns = 1; ns <= NEGF_SUB_NS; ++ns)
        if (y.fast_n[ns] > 0) return false;
    return true;
}

// Transfer, scan and sweeps of one layout on a stream.  With ev0/ev1 (only
// for a single fused launch) the events are recorded by that launch itself
// (hipExtLaunchKernel), so their interval is the kernel's own.
static int launch_rgf_kernels(const RgfLayout& y, const WLaunch& lp, hipStream_t st, hipEvent_t ev0 = nullptr,
                              hipEvent_t ev1 = nullptr)
{
    if (ev0 && single_fused_launch(y)) {
        hipExtLaunchKernelGGL((rgf_eq_fused_kernel<NEGF_SUB_K, NEGF_SUB_NS>), dim3(2 * y.nsblk),
                              dim3(64 * EQF_WAVES), 0, st, ev0, ev1, 0, lp, y.nsblk);
        NEGF_HIP_TRY(hipGetLastError());
        return NEGF_OK;
    }
    if (y.any_unfused) {
        hipLaunchKernelGGL((transfer_kernel<NEGF_SUB_K>), dim3((y.nitems + 3) / 4), dim3(256), 0, st, lp);
        hipLaunchKernelGGL(scan_kernel, dim3(2 * y.nsblk), dim3(64 * SCAN_WAVES), 0, st, lp, y.nsblk);
    }
    if (y.any_fused)
        hipLaunchKernelGGL(transfer_scan_kernel, dim3(2 * y.nsblk), dim3(64 * FUSED_WAVES), 0, st, lp, y.nsblk);
    if (y.any_fused2)
        hipLaunchKernelGGL((rgf_eq_fused_kernel<NEGF_SUB_K, NEGF_SUB_NS>), dim3(2 * y.nsblk), dim3(64 * EQF_WAVES), 0,
                           st, lp, y.nsblk);
    for (int ns = 1; ns <= NEGF_SUB_NS; ++ns) {
        if (y.fast_n[ns] == 0) continue;
        WLaunch lf = lp;
        lf.item_list = y.item_list + y.fast_off[ns];
        lf.nlist = y.fast_n[ns];
        if (y.kind == KIND_EQ)
            launch_sweep_eq_full(ns, (y.fast_n[ns] + 3) / 4, lf, st);
        else
            launch_sweep_neq_full(ns, (y.fast_n[ns] + 3) / 4, lf, st);
    }
    WLaunch ls = lp;
    ls.item_list = y.item_list + y.slow_off;
    ls.nlist = y.slow_n;
    const int nblk_slow = (y.slow_n + 3) / 4;
    if (y.slow_n == 0) {
        // nothing left for the general sweep
    } else if (y.kind == KIND_EQ) {
        if (y.write_G)
            launch_sweep<KIND_EQ, true>(nblk_slow, ls, st);
        else
            launch_sweep<KIND_EQ, false>(nblk_slow, ls, st);
    } else if (y.kind == KIND_NONEQ) {
        launch_sweep<KIND_NONEQ, false>(nblk_slow, ls, st);
    } else {
        launch_sweep<KIND_DOS, false>(nblk_slow, ls, st);
    }
    NEGF_HIP_TRY(hipGetLastError());
    return NEGF_OK;
}

// Per-chain DOS traces, or the fixed-order reduction of the group rows.
static int launch_rgf_finish(const RgfLayout& y, const WLaunch& lp, double* d_out, hipStream_t st,
                             const int* tube_active)
{
    if (y.kind == KIND_DOS) {
        hipLaunchKernelGGL(dos_finish_kernel, dim3((y.nchains + 255) / 256), dim3(256), 0, st, lp, y.ntubes);
        NEGF_HIP_TRY(hipGetLastError());
        return NEGF_OK;
    }
    if (y.max_rows <= 16) {
        dim3 g0((y.max_N + 255) / 256, y.ntubes);
        hipLaunchKernelGGL(reduce_rows_few_kernel, g0, dim3(256), 0, st, y.part, y.red_src_off, y.red_nrows,
                           y.red_N, d_out, y.red_dst_off, y.tubes, tube_active);
    } else {
        dim3 g1((y.max_N + 15) / 16, y.ntubes);
        hipLaunchKernelGGL(reduce_rows1_kernel, g1, dim3(256), 0, st, y.part, y.red_src_off, y.red_nrows, y.red_N,
                           d_out, y.red_dst_off, y.tubes, tube_active);
    }
    NEGF_HIP_TRY(hipGetLastError());
    return NEGF_OK;
}

int RgfPlan::enqueue(double* d_out, double2* d_diagG, double* d_chain_out, hipStream_t st, const int* tube_active,
                     double* d_site_out, bool finish)
{
    if (nitems_ == 0) {
        set_error("RgfPlan::enqueue: plan not built");
        return NEGF_EARG;
    }
    if (write_G_ && !d_diagG) {
        set_error("diag G output requested without a buffer");
        return NEGF_EARG;
    }
    if (kind_ == KIND_DOS && !d_chain_out) {
        set_error("DOS output buffer missing");
        return NEGF_EARG;
    }
    if (kind_ != KIND_DOS && !d_out) {
        set_error("output buffer missing");
        return NEGF_EARG;
    }
    RgfLayout y;
    y.kind = kind_;
    y.write_G = write_G_;
    y.nitems = nitems_;
    y.nchains = nchains_;
    y.ntubes = ntubes_;
    y.nsblk = nsblk_;
    y.max_N = max_N_;
    y.max_rows = max_rows_;
    y.chains = d_chains_;
    y.tubes = d_tubes_.as<WTube>();
    y.item_tube = d_item_tube_.as<int>();
    y.chain_tube = d_chain_tube_.as<int>();
    y.sblk_tube = d_sblk_tube_.as<int>();
    y.sblk_chain = d_sblk_chain_.as<int>();
    y.item_list = d_item_list_.as<int>();
    for (int ns = 0; ns <= NEGF_SUB_NS; ++ns) {
        y.fast_off[ns] = fast_off_[ns];
        y.fast_n[ns] = fast_n_[ns];
    }
    y.slow_off = slow_off_;
    y.slow_n = slow_n_;
    y.Tm = d_Tm_.as<double>();
    y.St = d_St_.as<double>();
    y.Dd = d_Dd_.as<double>();
    y.part = d_part_.as<double>();
    y.dos_part = d_dos_part_.as<double>();
    y.status = d_status_.as<int>();
    y.red_src_off = d_red_src_off_.as<long long>();
    y.red_nrows = d_red_nrows_.as<int>();
    y.red_N = d_red_N_.as<int>();
    y.red_dst_off = d_red_dst_off_.as<long long>();
    y.dec = dec_;
    y.any_fused = any_fused_;
    y.any_fused2 = any_fused2_;
    y.any_unfused = any_unfused_;
    if (ntubes_ == 1 && tube0_.size() == sizeof(WTube)) y.tube1 = reinterpret_cast<const WTube*>(tube0_.data());
    const WLaunch lp = launch_params(y, d_diagG, d_chain_out, tube_active, d_site_out);

    if (dec_.use) {
        hipLaunchKernelGGL(sigma_dec_kernel, dim3((nchains_ + 255) / 256), dim3(256), 0, st, lp, y.chain_tube,
                           nchains_);
        NEGF_HIP_TRY(hipGetLastError());
    }
    if (timing_) {
        if (ev_used_ == ev_start_.size()) {
            hipEvent_t a, b;
            NEGF_HIP_TRY(hipEventCreate(&a));
            NEGF_HIP_TRY(hipEventCreate(&b));
            ev_start_.push_back(a);
            ev_stop_.push_back(b);
        }
    }
    // a single fused launch records the events itself; otherwise they bracket the launches
    const bool ext = timing_ && single_fused_launch(y);
    if (timing_ && !ext) NEGF_HIP_TRY(hipEventRecord(ev_start_[ev_used_], st));
    int rc = ext ? launch_rgf_kernels(y, lp, st, ev_start_[ev_used_], ev_stop_[ev_used_])
                 : launch_rgf_kernels(y, lp, st);
    if (rc) return rc;
    if (timing_) {
        if (!ext) NEGF_HIP_TRY(hipEventRecord(ev_stop_[ev_used_], st));
        ++ev_used_;
    }
    if (!finish && kind_ != KIND_DOS) return NEGF_OK;
    return launch_rgf_finish(y, lp, d_out, st, tube_active);
}

int RgfPlan::check_status(hipStream_t st, int* status)
{
    int h = 0;
    NEGF_HIP_TRY(hipMemcpyAsync(&h, d_status_.p, sizeof(int), hipMemcpyDeviceToHost, st));
    NEGF_HIP_TRY(hipStreamSynchronize(st));
    if (h) NEGF_HIP_TRY(hipMemsetAsync(d_status_.p, 0, sizeof(int), st));
    *status = h;
    if (h & 2) {  // bit 1: set by rgf_eq_fused_kernel when a wait exceeded its bound
        set_error("fused RGF kernel: a wavefront's wait for the v walk exceeded its bound");
        return NEGF_EDEVICE;
    }
    return NEGF_OK;
}

// ---------------------------------------------------------------- Level 1
__global__ void pack_eq_chains(int nE, const double2* E, const double2* coef, double t_even, ChainDesc* out)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nE) return;
    ChainDesc c = {};
    c.Er = E[k].x;
    c.Ei = E[k].y;
    c.cr = coef ? coef[k].x : 0.0;
    c.ci = coef ? coef[k].y : 0.0;
    c.t_even = t_even;
    c.degen = 1.0;
    c.out_slot = k;
    out[k] = c;
}

__global__ void pack_noneq_chains(int nE, const double2* E, const double* cL, const double* cR, const double* iL,
                                  const double* iR, double t_even, ChainDesc* out)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nE) return;
    ChainDesc c = {};
    c.Er = E[k].x;
    c.Ei = E[k].y;
    c.cL = cL[k];
    c.cR = cR[k];
    c.iL = iL ? iL[k] : cL[k];
    c.iR = iR ? iR[k] : cR[k];
    c.t_even = t_even;
    c.degen = 1.0;
    c.out_slot = k;
    out[k] = c;
}

__global__ void split_dos_kernel(int nE, const double* pairs, double* dos, double* trans)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nE) return;
    dos[k] = pairs[2 * k];
    trans[k] = pairs[2 * k + 1];
}

struct L1Cache {
    std::mutex mu;
    DevBuf chains, pairs;
    RgfPlan plan;
};
// Never destroyed: device memory must not be freed from static destructors
// that may run after the HIP runtime has shut down.
static L1Cache& l1()
{
    static L1Cache* c = new L1Cache;
    return *c;
}

static int check_params(const negf_chain_params* p, const void* dU, int nE, const void* dE)
{
    if (!p || !dU || !dE || nE <= 0) {
        set_error("invalid arguments (null pointer or nE <= 0)");
        return NEGF_EARG;
    }
    if (p->N < 2 || p->t_even <= 0 || p->t_odd <= 0) {
        set_error("invalid chain parameters (N=%d, t_even=%g, t_odd=%g)", p->N, p->t_even, p->t_odd);
        return NEGF_EARG;
    }
    return NEGF_OK;
}

static RgfTube one_tube(const negf_chain_params* p, const double* dU, int nE)
{
    RgfTube tb;
    tb.N = p->N;
    tb.dU = dU;
    tb.Uc0 = p->Uc[0];
    tb.Uc1 = p->Uc[1];
    tb.t_odd = p->t_odd;
    tb.chain0 = 0;
    tb.nchain = nE;
    tb.out_off = 0;
    tb.L = p->segment_length;
    // else-if order of the reference (NEGF_Common.cpp:3457-3495): a site
    // that matches an earlier sample is not sampled again
    tb.sample[0] = p->N / 2;
    tb.sample[1] = p->N / 4;
    tb.sample[2] = p->N * 3 / 4;
    if (tb.sample[1] == tb.sample[0]) tb.sample[1] = -1;
    if (tb.sample[2] == tb.sample[0] || tb.sample[2] == tb.sample[1]) tb.sample[2] = -1;
    return tb;
}

static DecParams dec_of(const negf_chain_params* p)
{
    DecParams d;
    d.use = p->use_decimation;
    d.layers = p->decimation_layers;
    d.max_iter = p->decimation_max_iter;
    d.rel_error = p->decimation_rel_error;
    return d;
}

// Synchronise; report rootless surface GFs (g = 0 used) only if asked.
static int finish_status(RgfPlan& plan, hipStream_t st, bool report_root = false)
{
    int status = 0;
    int rc = plan.check_status(st, &status);
    if (rc) return rc;
    if ((status & 1) && report_root) {
        set_error("surface Green's function: no root with Im<0 at some energy (g = 0 used)");
        return NEGF_EROOT;
    }
    return NEGF_OK;
}

// ---------------------------------------------------------------- fixed scratch
// Workspace of negf_rgf_diag_accumulate_noalloc: a file-scope device array of
// NEGF_SCRATCH_BYTES that every GPU the library runs on holds from module load
// to process end.  In host code the array's name is only a host-side handle:
// the device address of each GPU's copy comes from hipGetSymbolAddress, taken
// once per device with that device current, and only that address is passed
// to kernels.
#define NEGF_SCRATCH_BYTES (64ull << 20)
#define NEGF_SCRATCH_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[NEGF_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_addr[NEGF_SCRATCH_MAX_DEVICES];

static int scratch_address(void** out)
{
    int dev = 0;
    NEGF_HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= NEGF_SCRATCH_MAX_DEVICES) {
        set_error("device %d outside the %d devices the fixed scratch tracks", dev, NEGF_SCRATCH_MAX_DEVICES);
        return NEGF_ENOTSUP;
    }
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error("hipGetSymbolAddress(scratch) on device %d failed: %s", dev, hipGetErrorString(e));
            return NEGF_EDEVICE;
        }
        g_scratch_addr[dev] = ptr;
    }
    *out = g_scratch_addr[dev];
    return NEGF_OK;
}

// Tables of a one-tube launch, written on the device: item -> tube, the item
// list (full segments first when the unrolled sweep takes them, in the order
// RgfPlan::build gives), chain groups, the row-reduction tables, the tube
// descriptor, and the status word cleared.
__global__ void one_tube_tables_kernel(WTube tb, WTube* tubes, int items, int P, int nfull, int fast, int* item_tube,
                                       int* item_list, int G, int rows, int* sblk_tube, int* sblk_chain, int nE,
                                       int* chain_tube, long long* red_src_off, int* red_nrows, int* red_N,
                                       long long* red_dst_off, int* status)
{
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    const int nf = fast ? G * nfull : 0;
    for (int i = i0; i < items; i += stride) {
        item_tube[i] = 0;
        int item = i;
        if (fast) item = i < nf ? (i / nfull) * P + i % nfull : (i - nf) * P + (P - 1);
        item_list[i] = item;
    }
    for (int g = i0; g < G; g += stride) {
        sblk_tube[g] = 0;
        sblk_chain[g] = g;
    }
    for (int c = i0; c < nE; c += stride) chain_tube[c] = 0;
    if (i0 == 0) {
        *tubes = tb;
        red_src_off[0] = 0;
        red_nrows[0] = rows;
        red_N[0] = tb.N;
        red_dst_off[0] = 0;
        *status = 0;
    }
}

}  // namespace negf

using namespace negf;

extern "C" {

const char* negf_last_error(void) { return g_err.c_str(); }
int negf_version(void) { return 100; }
int negf_max_sites(void) { return max_sites(); }

int negf_rgf_diag_accumulate(const negf_chain_params* p, const double* d_U, int nE, const negf_complex* d_E,
                             const negf_complex* d_coef, double* d_acc, negf_complex* d_diagG, void* stream)
{
    clear_error();
    int rc = check_params(p, d_U, nE, d_E);
    if (rc) return rc;
    if (!d_acc) {
        set_error("d_acc is NULL");
        return NEGF_EARG;
    }
    hipStream_t st = (hipStream_t)stream;
    L1Cache& c = l1();
    std::lock_guard<std::mutex> g(c.mu);
    if ((rc = c.chains.ensure(sizeof(ChainDesc) * nE))) return rc;
    hipLaunchKernelGGL(pack_eq_chains, dim3((nE + 255) / 256), dim3(256), 0, st, nE, (const double2*)d_E,
                       (const double2*)d_coef, p->t_even, c.chains.as<ChainDesc>());
    NEGF_HIP_TRY(hipGetLastError());
    std::vector<RgfTube> tubes{one_tube(p, d_U, nE)};
    if ((rc = c.plan.build(KIND_EQ, d_diagG != nullptr, tubes, c.chains.as<ChainDesc>(), nE, dec_of(p)))) return rc;
    if ((rc = c.plan.enqueue(d_acc, (double2*)d_diagG, nullptr, st))) return rc;
    return finish_status(c.plan, st);
}

int negf_rgf_diag_accumulate_noalloc(const negf_chain_params* p, const double* d_U, int nE,
                                     const negf_complex* d_E, const negf_complex* d_coef, double* d_acc,
                                     negf_complex* d_diagG, void* stream)
{
    clear_error();
    int rc = check_params(p, d_U, nE, d_E);
    if (rc) return rc;
    if (!d_acc) {
        set_error("d_acc is NULL");
        return NEGF_EARG;
    }
    if (p->use_decimation && (p->decimation_layers < 2 || p->decimation_layers > NEGF_MAX_DECIMATION_LAYERS)) {
        set_error("decimation_layers must be in [2, %d]", NEGF_MAX_DECIMATION_LAYERS);
        return NEGF_EARG;
    }
    int P, L;
    if (!choose_segments(p->N, p->segment_length, &P, &L)) {
        set_error("chain of %d sites exceeds the supported maximum %d", p->N, max_sites());
        return NEGF_ENOTSUP;
    }
    const size_t N = (size_t)p->N, E = (size_t)nE;
    const size_t G = (E + 63) / 64, items = G * (size_t)P;
    // the same launch forms as RgfPlan::build
    WTube tb = make_wtube(one_tube(p, d_U, nE), P, L, 0, 0);
    if (tb.fused && !d_diagG && L == NEGF_SEG_MAX && p->N % L == 0) tb.fused = 2;
    const size_t rows = tb.fused == 2 ? 2 * G : G;
    size_t off = 0;
    auto take = [&off](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_ch = take(E * sizeof(ChainDesc)), o_tb = take(sizeof(WTube));
    const size_t o_it = take(items * sizeof(int)), o_il = take(items * sizeof(int));
    const size_t o_ct = take(E * sizeof(int)), o_gt = take(G * sizeof(int)), o_gc = take(G * sizeof(int));
    const size_t o_Tm = take(items * 9 * 64 * sizeof(double)), o_St = take(items * 10 * 64 * sizeof(double));
    const size_t o_Dd = take(E * 3 * sizeof(double)), o_pt = take(rows * N * sizeof(double));
    const size_t o_r1 = take(sizeof(long long)), o_r2 = take(sizeof(int)), o_r3 = take(sizeof(int));
    const size_t o_r4 = take(sizeof(long long)), o_stat = take(sizeof(int));
    if (off > NEGF_SCRATCH_BYTES) {
        set_error("%d energies x %d sites need %zu bytes of workspace, more than the %llu-byte fixed scratch of "
                  "negf_rgf_diag_accumulate_noalloc (use negf_rgf_diag_accumulate)",
                  nE, p->N, off, (unsigned long long)NEGF_SCRATCH_BYTES);
        return NEGF_EARG;
    }
    void* base = nullptr;
    if ((rc = scratch_address(&base))) return rc;
    unsigned char* b = static_cast<unsigned char*>(base);
    hipStream_t st = (hipStream_t)stream;

    RgfLayout y;
    y.kind = KIND_EQ;
    y.write_G = d_diagG != nullptr;
    y.nitems = (int)items;
    y.nchains = nE;
    y.ntubes = 1;
    y.nsblk = (int)G;
    y.max_N = p->N;
    y.max_rows = (int)rows;
    y.chains = reinterpret_cast<ChainDesc*>(b + o_ch);
    WTube* d_tubes = reinterpret_cast<WTube*>(b + o_tb);
    y.tubes = d_tubes;
    y.item_tube = reinterpret_cast<int*>(b + o_it);
    y.item_list = reinterpret_cast<int*>(b + o_il);
    y.chain_tube = reinterpret_cast<int*>(b + o_ct);
    y.sblk_tube = reinterpret_cast<int*>(b + o_gt);
    y.sblk_chain = reinterpret_cast<int*>(b + o_gc);
    y.Tm = reinterpret_cast<double*>(b + o_Tm);
    y.St = reinterpret_cast<double*>(b + o_St);
    y.Dd = reinterpret_cast<double*>(b + o_Dd);
    y.part = reinterpret_cast<double*>(b + o_pt);
    y.red_src_off = reinterpret_cast<long long*>(b + o_r1);
    y.red_nrows = reinterpret_cast<int*>(b + o_r2);
    y.red_N = reinterpret_cast<int*>(b + o_r3);
    y.red_dst_off = reinterpret_cast<long long*>(b + o_r4);
    y.status = reinterpret_cast<int*>(b + o_stat);
    y.dec = dec_of(p);
    // the same split as RgfPlan::build: full segments to the unrolled sweep
    // unless diag G is written
    const int nfull = (p->N % L == 0) ? P : P - 1;
    const bool fast = !y.write_G;
    if (tb.fused == 2) {
        // no separate sweep
    } else if (fast) {
        y.fast_off[L / NEGF_SUB_K] = 0;
        y.fast_n[L / NEGF_SUB_K] = (int)G * nfull;
        y.slow_off = (int)G * nfull;
        y.slow_n = (int)items - (int)G * nfull;
    } else {
        y.slow_off = 0;
        y.slow_n = (int)items;
    }
    y.any_fused = tb.fused == 1;
    y.any_fused2 = tb.fused == 2;
    y.any_unfused = tb.fused == 0;
    y.tube1 = &tb;

    hipLaunchKernelGGL(pack_eq_chains, dim3((nE + 255) / 256), dim3(256), 0, st, nE, (const double2*)d_E,
                       (const double2*)d_coef, p->t_even, y.chains);
    const int nt = (int)std::min<size_t>(1024, (std::max(items, E) + 255) / 256);
    hipLaunchKernelGGL(one_tube_tables_kernel, dim3(nt), dim3(256), 0, st, tb, d_tubes, (int)items, P, nfull,
                       fast ? 1 : 0, const_cast<int*>(y.item_tube), const_cast<int*>(y.item_list), (int)G, (int)rows,
                       const_cast<int*>(y.sblk_tube), const_cast<int*>(y.sblk_chain), nE,
                       const_cast<int*>(y.chain_tube), const_cast<long long*>(y.red_src_off),
                       const_cast<int*>(y.red_nrows), const_cast<int*>(y.red_N),
                       const_cast<long long*>(y.red_dst_off), y.status);
    NEGF_HIP_TRY(hipGetLastError());
    const WLaunch lp = launch_params(y, (double2*)d_diagG, nullptr, nullptr);
    if (y.dec.use) {
        hipLaunchKernelGGL(sigma_dec_kernel, dim3((nE + 255) / 256), dim3(256), 0, st, lp, y.chain_tube, nE);
        NEGF_HIP_TRY(hipGetLastError());
    }
    if ((rc = launch_rgf_kernels(y, lp, st))) return rc;
    if ((rc = launch_rgf_finish(y, lp, d_acc, st, nullptr))) return rc;
    // like negf_rgf_diag_accumulate: the results are complete on return
    NEGF_HIP_TRY(hipStreamSynchronize(st));
    return NEGF_OK;
}

int negf_rgf_noneq_accumulate(const negf_chain_params* p, const double* d_U, int nE, const negf_complex* d_E,
                              const double* d_coefL, const double* d_coefR, const double* d_iL,
                              const double* d_iR, double* d_acc, double* d_integrand, void* stream)
{
    clear_error();
    int rc = check_params(p, d_U, nE, d_E);
    if (rc) return rc;
    if (!d_acc || !d_coefL || !d_coefR) {
        set_error("d_acc / d_coefL / d_coefR is NULL");
        return NEGF_EARG;
    }
    hipStream_t st = (hipStream_t)stream;
    L1Cache& c = l1();
    std::lock_guard<std::mutex> g(c.mu);
    if ((rc = c.chains.ensure(sizeof(ChainDesc) * nE))) return rc;
    hipLaunchKernelGGL(pack_noneq_chains, dim3((nE + 255) / 256), dim3(256), 0, st, nE, (const double2*)d_E,
                       d_coefL, d_coefR, d_iL, d_iR, p->t_even, c.chains.as<ChainDesc>());
    NEGF_HIP_TRY(hipGetLastError());
    std::vector<RgfTube> tubes{one_tube(p, d_U, nE)};
    if (!d_integrand)
        for (int q = 0; q < 3; ++q) tubes[0].sample[q] = -1;
    if ((rc = c.plan.build(KIND_NONEQ, false, tubes, c.chains.as<ChainDesc>(), nE, dec_of(p)))) return rc;
    if ((rc = c.plan.enqueue(d_acc, nullptr, d_integrand, st))) return rc;
    return finish_status(c.plan, st);
}

int negf_rgf_dos(const negf_chain_params* p, const double* d_U, int nE, const negf_complex* d_E, double* d_dos,
                 double* d_trans, void* stream)
{
    clear_error();
    int rc = check_params(p, d_U, nE, d_E);
    if (rc) return rc;
    if (!d_dos || !d_trans) {
        set_error("d_dos / d_trans is NULL");
        return NEGF_EARG;
    }
    hipStream_t st = (hipStream_t)stream;
    L1Cache& c = l1();
    std::lock_guard<std::mutex> g(c.mu);
    if ((rc = c.chains.ensure(sizeof(ChainDesc) * nE))) return rc;
    if ((rc = c.pairs.ensure(sizeof(double) * 2 * nE))) return rc;
    hipLaunchKernelGGL(pack_eq_chains, dim3((nE + 255) / 256), dim3(256), 0, st, nE, (const double2*)d_E,
                       (const double2*)nullptr, p->t_even, c.chains.as<ChainDesc>());
    NEGF_HIP_TRY(hipGetLastError());
    std::vector<RgfTube> tubes{one_tube(p, d_U, nE)};
    if ((rc = c.plan.build(KIND_DOS, false, tubes, c.chains.as<ChainDesc>(), nE, dec_of(p)))) return rc;
    if ((rc = c.plan.enqueue(nullptr, nullptr, c.pairs.as<double>(), st))) return rc;
    hipLaunchKernelGGL(split_dos_kernel, dim3((nE + 255) / 256), dim3(256), 0, st, nE, c.pairs.as<double>(), d_dos,
                       d_trans);
    NEGF_HIP_TRY(hipGetLastError());
    return finish_status(c.plan, st);
}

struct negf_rgf_plan {
    negf::RgfPlan plan;
    negf::DevBuf chains;
    int nE = 0, N = 0;
    bool want_G = false;
};

int negf_rgf_plan_create(const negf_chain_params* p, const double* d_U, int nE, const negf_complex* d_E,
                         const negf_complex* d_coef, int want_diagG, negf_rgf_plan** out)
{
    clear_error();
    if (!out) {
        set_error("out is NULL");
        return NEGF_EARG;
    }
    int rc = check_params(p, d_U, nE, d_E);
    if (rc) return rc;
    negf_rgf_plan* h = new negf_rgf_plan;
    h->nE = nE;
    h->N = p->N;
    h->want_G = want_diagG != 0;
    if ((rc = h->chains.ensure(sizeof(ChainDesc) * nE))) {
        delete h;
        return rc;
    }
    hipLaunchKernelGGL(pack_eq_chains, dim3((nE + 255) / 256), dim3(256), 0, 0, nE, (const double2*)d_E,
                       (const double2*)d_coef, p->t_even, h->chains.as<ChainDesc>());
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        set_error("pack_eq_chains failed: %s", hipGetErrorString(e));
        delete h;
        return NEGF_EDEVICE;
    }
    std::vector<RgfTube> tubes{one_tube(p, d_U, nE)};
    if ((rc = h->plan.build(KIND_EQ, h->want_G, tubes, h->chains.as<ChainDesc>(), nE, dec_of(p)))) {
        delete h;
        return rc;
    }
    *out = h;
    return NEGF_OK;
}

int negf_rgf_plan_execute(negf_rgf_plan* h, double* d_acc, negf_complex* d_diagG, void* stream)
{
    clear_error();
    if (!h || !d_acc || (h->want_G && !d_diagG)) {
        set_error("invalid plan execute arguments");
        return NEGF_EARG;
    }
    return h->plan.enqueue(d_acc, (double2*)d_diagG, nullptr, (hipStream_t)stream);
}

int negf_rgf_plan_status(negf_rgf_plan* h, void* stream)
{
    clear_error();
    if (!h) {
        set_error("plan is NULL");
        return NEGF_EARG;
    }
    return finish_status(h->plan, (hipStream_t)stream, true);
}

int negf_rgf_plan_enable_timing(negf_rgf_plan* h, int on)
{
    if (!h) return NEGF_EARG;
    h->plan.enable_timing(on != 0);
    return NEGF_OK;
}

int negf_rgf_plan_kernel_time(negf_rgf_plan* h, double* total_ms, int* launches)
{
    clear_error();
    if (!h || !total_ms || !launches) {
        set_error("invalid arguments");
        return NEGF_EARG;
    }
    return h->plan.kernel_time(total_ms, launches);
}

int negf_rgf_plan_destroy(negf_rgf_plan* h)
{
    if (!h) return NEGF_EARG;
    (void)hipDeviceSynchronize();
    delete h;
    return NEGF_OK;
}

int negf_gl_path(negf_complex a, negf_complex b, int n, int type, negf_complex* E, double* w, negf_complex* mul)
{
    clear_error();
    if (n <= 0 || !E || !w || !mul || (type != 0 && type != 1)) {
        set_error("invalid gl path arguments");
        return NEGF_EARG;
    }
    gl_path(a.re, a.im, b.re, b.im, n, type, (double*)E, w, (double*)mul);
    return NEGF_OK;
}

}  // extern "C"
