// negf_engine.hip -- host launcher for the segmented RGF kernels and the
// Level-1 C ABI (batched RGF over an energy list for one chain).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "negf_internal.h"
#include "rgf_seg.hip"

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

// ---------------------------------------------------------------- variants
struct VariantInfo {
    int K, NS;
};
static const VariantInfo kVariants[] = {{16, 4}, {32, 8}};
static const int kNumVariants = 2;

int choose_variant(int N, int* variant, int* Lf)
{
    int L0 = (N + NEGF_SEGMENTS - 1) / NEGF_SEGMENTS;
    for (int v = 0; v < kNumVariants; ++v) {
        int K = kVariants[v].K;
        int L = (L0 + K - 1) / K * K;
        if (L <= K * kVariants[v].NS) {
            *variant = v;
            *Lf = L;
            return 1;
        }
    }
    return 0;
}
int max_sites() { return NEGF_SEGMENTS * kVariants[kNumVariants - 1].K * kVariants[kNumVariants - 1].NS; }

template <int K, int NS>
static hipError_t launch_kernel(int kind, bool wg, int nblocks, int lds, const LaunchParams& lp, hipStream_t st)
{
    dim3 grid(nblocks), block(NEGF_BLOCK_THREADS);
    if (kind == KIND_EQ) {
        if (wg)
            hipLaunchKernelGGL((rgf_seg_kernel<K, NS, KIND_EQ, true>), grid, block, lds, st, lp);
        else
            hipLaunchKernelGGL((rgf_seg_kernel<K, NS, KIND_EQ, false>), grid, block, lds, st, lp);
    } else if (kind == KIND_NONEQ) {
        hipLaunchKernelGGL((rgf_seg_kernel<K, NS, KIND_NONEQ, false>), grid, block, lds, st, lp);
    } else {
        hipLaunchKernelGGL((rgf_seg_kernel<K, NS, KIND_DOS, false>), grid, block, 0, st, lp);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- plan
RgfPlan::~RgfPlan()
{
    DevBuf* bufs[] = {&d_tubes_,   &d_bchain0_,       &d_bnchain_,    &d_btube_,       &d_bpart_,
                      &d_Ut_,      &d_part_,          &d_tmp_,        &d_status_,      &d_red_src_off_,
                      &d_red_nrows_, &d_red_N_,       &d_red_dst_off_, &d_red2_src_off_, &d_red2_nrows_,
                      &d_red2_dst_off_};
    for (DevBuf* b : bufs) b->release();
}

template <class T>
static int upload(DevBuf& b, const std::vector<T>& v)
{
    int rc = b.ensure(std::max<size_t>(1, v.size()) * sizeof(T));
    if (rc) return rc;
    if (!v.empty()) NEGF_HIP_TRY(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return NEGF_OK;
}

int RgfPlan::build(int kind, bool write_G, const std::vector<RgfTube>& tubes, const ChainDesc* d_chains,
                   const DecParams& dec)
{
    if (tubes.empty()) {
        set_error("RgfPlan::build: no tubes");
        return NEGF_EARG;
    }
    if (dec.use && (dec.layers < 2 || dec.layers > NEGF_MAX_DECIMATION_LAYERS)) {
        set_error("decimation_layers must be in [2, %d]", NEGF_MAX_DECIMATION_LAYERS);
        return NEGF_EARG;
    }
    kind_ = kind;
    write_G_ = write_G;
    d_chains_ = d_chains;
    dec_ = dec;
    tubes_ = tubes;
    ntubes_ = (int)tubes.size();
    Lf_.assign(ntubes_, 0);
    u_off_.assign(ntubes_, 0);
    variant_ = -1;
    long long uoff = 0;
    max_N_ = 0;
    lds_bytes_ = 0;
    any_global_atomic_ = false;
    for (int t = 0; t < ntubes_; ++t) {
        const RgfTube& tb = tubes[t];
        if (tb.N < 2 || !tb.dU || tb.nchain < 0 || tb.t_odd <= 0) {
            set_error("RgfPlan::build: invalid tube %d (N=%d)", t, tb.N);
            return NEGF_EARG;
        }
        int v, L;
        if (!choose_variant(tb.N, &v, &L)) {
            set_error("chain of %d sites exceeds the supported maximum %d", tb.N, max_sites());
            return NEGF_ENOTSUP;
        }
        variant_ = std::max(variant_, v);
        Lf_[t] = L;
        max_N_ = std::max(max_N_, tb.N);
    }
    // one variant per launch: recompute Lf for the common (largest) variant
    for (int t = 0; t < ntubes_; ++t) {
        int K = kVariants[variant_].K;
        int L0 = (tubes[t].N + NEGF_SEGMENTS - 1) / NEGF_SEGMENTS;
        Lf_[t] = (L0 + K - 1) / K * K;
        u_off_[t] = uoff;
        uoff += (long long)Lf_[t] * NEGF_SEGMENTS;
        if (tubes[t].N <= NEGF_LDS_SITES_MAX)
            lds_bytes_ = std::max(lds_bytes_, tubes[t].N * (int)sizeof(double));
        else
            any_global_atomic_ = true;
    }
    int rc = d_Ut_.ensure(uoff * sizeof(double));
    if (rc) return rc;

    // tube descriptors
    std::vector<TubeDesc> td(ntubes_);
    for (int t = 0; t < ntubes_; ++t) {
        const RgfTube& tb = tubes[t];
        TubeDesc& d = td[t];
        std::memset(&d, 0, sizeof(d));
        d.N = tb.N;
        d.seg_len = Lf_[t];
        d.nsub = Lf_[t] / kVariants[variant_].K;
        d.lds_ok = tb.N <= NEGF_LDS_SITES_MAX;
        d.t_odd = tb.t_odd;
        d.Uc0 = tb.Uc0;
        d.Uc1 = tb.Uc1;
        d.u_off = u_off_[t];
        d.out_off = tb.out_off;
        for (int q = 0; q < 3; ++q) d.sample[q] = tb.sample[q];
    }
    // blocks
    std::vector<int> bchain0, bnchain, btube;
    std::vector<long long> bpart;
    std::vector<long long> rsrc, rdst;
    std::vector<int> rnrows, rN;
    long long poff = 0;
    max_rows_ = 0;
    for (int t = 0; t < ntubes_; ++t) {
        const RgfTube& tb = tubes[t];
        int nb = (tb.nchain + NEGF_CHAINS_PER_BLOCK - 1) / NEGF_CHAINS_PER_BLOCK;
        rsrc.push_back(poff);
        rnrows.push_back(nb);
        rN.push_back(tb.N);
        for (int b = 0; b < nb; ++b) {
            bchain0.push_back(tb.chain0 + b * NEGF_CHAINS_PER_BLOCK);
            bnchain.push_back(std::min(NEGF_CHAINS_PER_BLOCK, tb.nchain - b * NEGF_CHAINS_PER_BLOCK));
            btube.push_back(t);
            bpart.push_back(poff + (long long)b * tb.N);
        }
        poff += (long long)nb * tb.N;
        max_rows_ = std::max(max_rows_, nb);
    }
    nblocks_ = (int)bchain0.size();
    part_elems_ = poff;
    if ((rc = d_part_.ensure(std::max<long long>(1, poff) * sizeof(double)))) return rc;
    // reduction pass 1: chunks of reduce_chunk_ rows -> tmp; pass 2: tmp -> out
    reduce_chunk_ = 16;
    max_chunks_ = (max_rows_ + reduce_chunk_ - 1) / reduce_chunk_;
    long long toff = 0;
    std::vector<long long> tdst, t2src;
    std::vector<int> t2rows;
    for (int t = 0; t < ntubes_; ++t) {
        int nch = (rnrows[t] + reduce_chunk_ - 1) / reduce_chunk_;
        tdst.push_back(toff);
        t2src.push_back(toff);
        t2rows.push_back(nch);
        toff += (long long)nch * rN[t];
        rdst.push_back(tubes[t].out_off);
    }
    tmp_elems_ = toff;
    if ((rc = d_tmp_.ensure(std::max<long long>(1, toff) * sizeof(double)))) return rc;
    if ((rc = upload(d_tubes_, td))) return rc;
    if ((rc = upload(d_bchain0_, bchain0))) return rc;
    if ((rc = upload(d_bnchain_, bnchain))) return rc;
    if ((rc = upload(d_btube_, btube))) return rc;
    if ((rc = upload(d_bpart_, bpart))) return rc;
    if ((rc = upload(d_red_src_off_, rsrc))) return rc;
    if ((rc = upload(d_red_nrows_, rnrows))) return rc;
    if ((rc = upload(d_red_N_, rN))) return rc;
    if ((rc = upload(d_red_dst_off_, tdst))) return rc;
    if ((rc = upload(d_red2_src_off_, t2src))) return rc;
    if ((rc = upload(d_red2_nrows_, t2rows))) return rc;
    if ((rc = upload(d_red2_dst_off_, rdst))) return rc;
    if ((rc = d_status_.ensure(sizeof(int)))) return rc;
    NEGF_HIP_TRY(hipMemset(d_status_.p, 0, sizeof(int)));
    return NEGF_OK;
}

int RgfPlan::enqueue(double* d_out, double2* d_diagG, double* d_chain_out, hipStream_t st)
{
    if (variant_ < 0 || nblocks_ == 0) {
        set_error("RgfPlan::enqueue: plan not built");
        return NEGF_EARG;
    }
    if (write_G_ && !d_diagG) {
        set_error("diag G output requested without a buffer");
        return NEGF_EARG;
    }
    if ((kind_ == KIND_DOS) && !d_chain_out) {
        set_error("DOS output buffer missing");
        return NEGF_EARG;
    }
    // transposed potentials
    for (int t = 0; t < ntubes_; ++t) {
        int total = Lf_[t] * NEGF_SEGMENTS;
        hipLaunchKernelGGL(transpose_u_kernel, dim3((total + 255) / 256), dim3(256), 0, st, tubes_[t].dU,
                           tubes_[t].N, Lf_[t], d_Ut_.as<double>() + u_off_[t]);
    }
    if (any_global_atomic_ && kind_ != KIND_DOS) {
        hipLaunchKernelGGL(fill_kernel, dim3((part_elems_ + 255) / 256), dim3(256), 0, st, d_part_.as<double>(),
                           part_elems_, 0.0);
    }
    NEGF_HIP_TRY(hipGetLastError());
    LaunchParams lp;
    lp.chains = const_cast<ChainDesc*>(d_chains_);
    lp.tubes = d_tubes_.as<TubeDesc>();
    lp.block_chain0 = d_bchain0_.as<int>();
    lp.block_nchain = d_bnchain_.as<int>();
    lp.block_tube = d_btube_.as<int>();
    lp.block_part = d_bpart_.as<long long>();
    lp.Ut = d_Ut_.as<double>();
    lp.part = d_part_.as<double>();
    lp.diagG = d_diagG;
    lp.chain_out = d_chain_out;
    lp.status = d_status_.as<int>();
    lp.use_decimation = dec_.use;
    lp.dec_layers = dec_.layers;
    lp.dec_max_iter = dec_.max_iter;
    lp.dec_rel_error = dec_.rel_error;
    hipLaunchKernelGGL(sigma_kernel, dim3(nblocks_), dim3(64), 0, st, lp);
    NEGF_HIP_TRY(hipGetLastError());
    hipError_t e;
    if (variant_ == 0)
        e = launch_kernel<16, 4>(kind_, write_G_, nblocks_, lds_bytes_, lp, st);
    else
        e = launch_kernel<32, 8>(kind_, write_G_, nblocks_, lds_bytes_, lp, st);
    if (e != hipSuccess) {
        set_error("rgf kernel launch failed: %s", hipGetErrorString(e));
        return NEGF_EDEVICE;
    }
    if (kind_ == KIND_DOS) return NEGF_OK;
    if (!d_out) {
        set_error("output buffer missing");
        return NEGF_EARG;
    }
    // fixed-order reduction of the partial rows
    {
        dim3 g1((max_N_ + 255) / 256, max_chunks_, ntubes_);
        hipLaunchKernelGGL(reduce_rows_kernel, g1, dim3(256), 0, st, d_part_.as<double>(),
                           d_red_src_off_.as<long long>(), d_red_nrows_.as<int>(), d_red_N_.as<int>(),
                           reduce_chunk_, d_tmp_.as<double>(), d_red_dst_off_.as<long long>());
        dim3 g2((max_N_ + 255) / 256, 1, ntubes_);
        hipLaunchKernelGGL(reduce_rows_kernel, g2, dim3(256), 0, st, d_tmp_.as<double>(),
                           d_red2_src_off_.as<long long>(), d_red2_nrows_.as<int>(), d_red_N_.as<int>(),
                           1 << 30, d_out, d_red2_dst_off_.as<long long>());
    }
    NEGF_HIP_TRY(hipGetLastError());
    return NEGF_OK;
}

int RgfPlan::check_status(hipStream_t st, int* status)
{
    int h = 0;
    NEGF_HIP_TRY(hipMemcpyAsync(&h, d_status_.p, sizeof(int), hipMemcpyDeviceToHost, st));
    NEGF_HIP_TRY(hipStreamSynchronize(st));
    if (h) NEGF_HIP_TRY(hipMemsetAsync(d_status_.p, 0, sizeof(int), st));
    *status = h;
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
static L1Cache& l1() {
    static L1Cache c;
    return c;
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

static int finish_status(RgfPlan& plan, hipStream_t st)
{
    int status = 0;
    int rc = plan.check_status(st, &status);
    if (rc) return rc;
    if (status & 1) {
        set_error("surface Green's function: no root with Im<0 at some energy");
        return NEGF_EROOT;
    }
    return NEGF_OK;
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
    if ((rc = c.plan.build(KIND_EQ, d_diagG != nullptr, tubes, c.chains.as<ChainDesc>(), dec_of(p)))) return rc;
    if ((rc = c.plan.enqueue(d_acc, (double2*)d_diagG, nullptr, st))) return rc;
    return finish_status(c.plan, st);
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
    if ((rc = c.plan.build(KIND_NONEQ, false, tubes, c.chains.as<ChainDesc>(), dec_of(p)))) return rc;
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
    if ((rc = c.plan.build(KIND_DOS, false, tubes, c.chains.as<ChainDesc>(), dec_of(p)))) return rc;
    if ((rc = c.plan.enqueue(nullptr, nullptr, c.pairs.as<double>(), st))) return rc;
    hipLaunchKernelGGL(split_dos_kernel, dim3((nE + 255) / 256), dim3(256), 0, st, nE, c.pairs.as<double>(), d_dos,
                       d_trans);
    NEGF_HIP_TRY(hipGetLastError());
    return finish_status(c.plan, st);
}

}  // extern "C"
