// negf_batch.hip -- many tubes solved together.
//
// The equilibrium chains of all tubes (every mode x contour point and pole of
// every tube) form one RGF plan, the non-equilibrium chains of all biased
// tubes a second one, so a batched solve is one launch per pipeline stage
// (self-energies, segment transfer matrices, segment scan, sweeps, row
// reduction) over every tube, however uneven their lengths, plus one combine
// kernel over the concatenated sites.  Each tube keeps its own parameters,
// contacts, paths and potential; the batch only concatenates.  The enqueue
// sequence can be captured once into a hipGraph and replayed; it is
// re-captured when a tube's paths change.  An autotuner picks the segment
// length by timing the solve.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include <algorithm>
#include <cstring>
#include <vector>

#include "negf_internal.h"
#include "negf_material.h"

namespace negf {

__global__ void batch_combine_kernel(long long n, const double* __restrict__ eq, const double* __restrict__ neq,
                                     const double* __restrict__ rho0, double* __restrict__ Q)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) Q[i] = (eq[i] + neq[i]) - rho0[i];
}

class Batch {
   public:
    ~Batch()
    {
        drop_graph();
        for (hipStream_t* sp : {&stream, &stream_neq})
            if (*sp) {
                (void)hipStreamSynchronize(*sp);
                (void)hipStreamDestroy(*sp);
            }
        for (hipEvent_t e : {ev0, ev1, ev_in, ev_out, ev_fork, ev_join})
            if (e) (void)hipEventDestroy(e);
    }

    int init(Tube* const* t, int n)
    {
        tubes.assign(t, t + n);
        if (n <= 0) {
            set_error("empty batch");
            return NEGF_EARG;
        }
        const DecParams d0 = tubes[0]->dec();
        for (Tube* tb : tubes) {
            const DecParams d = tb->dec();
            if (d.use != d0.use || d.layers != d0.layers || d.max_iter != d0.max_iter ||
                d.rel_error != d0.rel_error) {
                set_error("tubes of a batch must share the lead (decimation) settings");
                return NEGF_EARG;
            }
        }
        off.assign(n + 1, 0);
        for (int i = 0; i < n; ++i) off[i + 1] = off[i] + tubes[i]->N;
        seg_len = tubes[0]->p.segment_length;
        NEGF_HIP_TRY(hipStreamCreate(&stream));
        NEGF_HIP_TRY(hipEventCreate(&ev0));
        NEGF_HIP_TRY(hipEventCreate(&ev1));
        NEGF_HIP_TRY(hipEventCreateWithFlags(&ev_in, hipEventDisableTiming));
        NEGF_HIP_TRY(hipEventCreateWithFlags(&ev_out, hipEventDisableTiming));
        NEGF_HIP_TRY(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        NEGF_HIP_TRY(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
        NEGF_HIP_TRY(hipStreamCreate(&stream_neq));
        int rc;
        const long long Nt = off[n];
        if ((rc = d_rho0.ensure(sizeof(double) * Nt))) return rc;
        if ((rc = d_eq.ensure(sizeof(double) * Nt))) return rc;
        if ((rc = d_neq.ensure(sizeof(double) * Nt))) return rc;
        for (int i = 0; i < n; ++i) {
            NEGF_HIP_TRY(hipStreamSynchronize(tubes[i]->stream));
            NEGF_HIP_TRY(hipMemcpy(d_rho0.as<double>() + off[i], tubes[i]->d_rho0.p, sizeof(double) * tubes[i]->N,
                                   hipMemcpyDeviceToDevice));
        }
        return build();
    }

    // Concatenate the chains of every tube and build the two plans.
    int build()
    {
        drop_graph();
        // every rebuild may move the plans' buffers: graphs captured by
        // callers (the batched SCF) compare this counter before a replay
        ++generation;
        const int n = (int)tubes.size();
        std::vector<ChainDesc> eq, neq;
        std::vector<RgfTube> teq, tneq;
        for (int i = 0; i < n; ++i) {
            Tube& t = *tubes[i];
            RgfTube r = t.rgf_tube(t.d_U.as<double>(), t.d_Uc.as<double>(), (int)t.eq_host.size(), seg_len);
            r.chain0 = (int)eq.size();
            r.out_off = off[i];
            r.ext = i;
            eq.insert(eq.end(), t.eq_host.begin(), t.eq_host.end());
            teq.push_back(r);
            if (!t.neq_host.empty()) {
                RgfTube q = t.rgf_tube(t.d_U.as<double>(), t.d_Uc.as<double>(), (int)t.neq_host.size(), seg_len);
                q.chain0 = (int)neq.size();
                q.out_off = off[i];
                q.ext = i;
                for (int w = 0; w < 3; ++w) q.sample[w] = -1;  // no integrand in the batched launch
                neq.insert(neq.end(), t.neq_host.begin(), t.neq_host.end());
                tneq.push_back(q);
            }
        }
        int rc;
        if ((rc = upload(eq_chains, eq))) return rc;
        if ((rc = eq_plan.build(KIND_EQ, false, teq, eq_chains.as<ChainDesc>(), (int)eq.size(), tubes[0]->dec())))
            return rc;
        n_eq = (int)eq.size();
        n_neq = (int)neq.size();
        if (n_neq > 0) {
            if ((rc = upload(neq_chains, neq))) return rc;
            if ((rc = neq_plan.build(KIND_NONEQ, false, tneq, neq_chains.as<ChainDesc>(), n_neq,
                                     tubes[0]->dec())))
                return rc;
        }
        // tubes without a non-equilibrium part keep RhoNonEq = 0
        if ((rc = fill_device(d_neq.as<double>(), off[n], 0.0, stream))) return rc;
        NEGF_HIP_TRY(hipStreamSynchronize(stream));
        versions.resize(n);
        for (int i = 0; i < n; ++i) versions[i] = tubes[i]->paths_version;
        energy_sites = 0;
        for (int i = 0; i < n; ++i)
            energy_sites += (double)tubes[i]->N * (double)(tubes[i]->eq_host.size() + tubes[i]->neq_host.size());
        return NEGF_OK;
    }

    int enqueue(double* dQ) { return enqueue_on(dQ, stream); }
    // The equilibrium and non-equilibrium plans are independent: the second
    // runs on its own stream (forked from st and joined back before the
    // combine, a form that stream capture records too), so the latency-bound
    // walks of one overlap the other's sweeps.
    int enqueue_on(double* dQ, hipStream_t st, const int* tube_active = nullptr)
    {
        int rc;
        if (n_neq > 0) {
            NEGF_HIP_TRY(hipEventRecord(ev_fork, st));
            NEGF_HIP_TRY(hipStreamWaitEvent(stream_neq, ev_fork, 0));
            if ((rc = neq_plan.enqueue(d_neq.as<double>(), nullptr, nullptr, stream_neq, tube_active))) return rc;
            NEGF_HIP_TRY(hipEventRecord(ev_join, stream_neq));
        }
        if ((rc = eq_plan.enqueue(d_eq.as<double>(), nullptr, nullptr, st, tube_active))) return rc;
        if (n_neq > 0) NEGF_HIP_TRY(hipStreamWaitEvent(st, ev_join, 0));
        const long long Nt = off.back();
        hipLaunchKernelGGL(batch_combine_kernel, dim3((unsigned)((Nt + 255) / 256)), dim3(256), 0, st, Nt,
                           d_eq.as<double>(), d_neq.as<double>(), d_rho0.as<double>(), dQ);
        NEGF_HIP_TRY(hipGetLastError());
        return NEGF_OK;
    }
    // host part of a solve for every tube; rebuild the plans if any paths changed
    int prepare(int iter, bool* rebuilt)
    {
        int rc;
        bool changed = false;
        for (size_t i = 0; i < tubes.size(); ++i) {
            if ((rc = tubes[i]->prepare(iter))) return rc;
            changed |= tubes[i]->paths_version != versions[i];
        }
        if (changed && (rc = build())) return rc;
        *rebuilt = changed;
        return NEGF_OK;
    }

    int solve(int iter, double* dQ, hipStream_t st)
    {
        int rc;
        // host part per tube (contacts changed, adaptive corrections)
        bool changed = false;
        for (size_t i = 0; i < tubes.size(); ++i) {
            if ((rc = tubes[i]->prepare(iter))) return rc;
            changed |= tubes[i]->paths_version != versions[i];
        }
        if (changed && (rc = build())) return rc;
        // the batch stream follows the caller's stream and any single-tube
        // solve still in flight on a tube's own stream
        NEGF_HIP_TRY(hipEventRecord(ev_in, st));
        NEGF_HIP_TRY(hipStreamWaitEvent(stream, ev_in, 0));
        for (Tube* t : tubes) {
            if (!t->stream_pending) continue;
            NEGF_HIP_TRY(hipEventRecord(ev_in, t->stream));
            NEGF_HIP_TRY(hipStreamWaitEvent(stream, ev_in, 0));
            t->stream_pending = false;
        }
        if (timing) NEGF_HIP_TRY(hipEventRecord(ev0, stream));
        if (use_graph) {
            if (!graph_ok || graph_dQ != dQ) {
                drop_graph();
                NEGF_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                int erc = enqueue(dQ);
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamEndCapture(stream, &g);
                if (erc) {
                    if (g) (void)hipGraphDestroy(g);
                    return erc;
                }
                if (e != hipSuccess) {
                    set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
                    return NEGF_EDEVICE;
                }
                e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e != hipSuccess) {
                    set_error("hipGraphInstantiate: %s", hipGetErrorString(e));
                    gexec = nullptr;
                    return NEGF_EDEVICE;
                }
                graph_ok = true;
                graph_dQ = dQ;
            }
            NEGF_HIP_TRY(hipGraphLaunch(gexec, stream));
        } else {
            if ((rc = enqueue(dQ))) return rc;
        }
        if (timing) NEGF_HIP_TRY(hipEventRecord(ev1, stream));
        NEGF_HIP_TRY(hipEventRecord(ev_out, stream));
        NEGF_HIP_TRY(hipStreamWaitEvent(st, ev_out, 0));
        return NEGF_OK;
    }

    int last_ms(double* ms)
    {
        NEGF_HIP_TRY(hipEventSynchronize(ev1));
        float f = 0.f;
        NEGF_HIP_TRY(hipEventElapsedTime(&f, ev0, ev1));
        *ms = f;
        return NEGF_OK;
    }

    // Time the solve for each candidate segment length; keep the fastest.
    int autotune(double* dQ, int* best_L, double* best_ms)
    {
        static const int cand[] = {32, 48, 64, 96, 128};
        const int keep_L = seg_len;
        double best = 1e300;
        int bl = keep_L;
        int rc;
        for (int L : cand) {
            seg_len = L;
            if ((rc = build())) return rc;
            double tot = 0.0;
            for (int r = 0; r < 4; ++r) {
                NEGF_HIP_TRY(hipEventRecord(ev0, stream));
                if ((rc = enqueue(dQ))) return rc;
                NEGF_HIP_TRY(hipEventRecord(ev1, stream));
                double ms;
                if ((rc = last_ms(&ms))) return rc;
                if (r > 0) tot += ms;  // first run warms caches and code
            }
            tot /= 3.0;
            if (tot < best) {
                best = tot;
                bl = L;
            }
        }
        seg_len = bl;
        if ((rc = build())) return rc;
        if (best_L) *best_L = bl;
        if (best_ms) *best_ms = best;
        return NEGF_OK;
    }

    // Workload key of the tuning cache: device (architecture and compute
    // units) and, per tube, its sites and chain counts (FNV-1a hash).
    int tune_key(std::string& key)
    {
        int dev = 0;
        NEGF_HIP_TRY(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        NEGF_HIP_TRY(hipGetDeviceProperties(&prop, dev));
        unsigned long long h = 1469598103934665603ull;
        auto mix = [&h](long long v) {
            for (int b = 0; b < 8; ++b) {
                h ^= (unsigned long long)((v >> (8 * b)) & 0xff);
                h *= 1099511628211ull;
            }
        };
        for (Tube* t : tubes) {
            mix(t->N);
            mix((long long)t->eq_host.size());
            mix((long long)t->neq_host.size());
        }
        char buf[256];
        std::snprintf(buf, sizeof(buf), "rgf-batch-v1 %s cu%d tubes%d %016llx", prop.gcnArchName,
                      prop.multiProcessorCount, (int)tubes.size(), h);
        for (char* c = buf; *c; ++c)
            if (*c == ':') *c = '_';
        key = buf;
        return NEGF_OK;
    }
    // autotune() with a persistent cache: a text file of "key L ms" lines.
    // A hit applies the stored segment length without timing; a miss tunes
    // and appends the result.
    int autotune_cached(double* dQ, const char* path, int* best_L, double* best_ms, int* hit)
    {
        std::string key;
        int rc;
        if ((rc = tune_key(key))) return rc;
        *hit = 0;
        if (FILE* f = std::fopen(path, "r")) {
            char line[512];
            const size_t kl = key.size();
            while (std::fgets(line, sizeof(line), f)) {
                int L = 0;
                double ms = 0.0;
                if (!std::strncmp(line, key.c_str(), kl) && line[kl] == ' ' &&
                    std::sscanf(line + kl, "%d %lf", &L, &ms) == 2 && L > 0) {
                    *hit = 1;  // the last matching line wins
                    *best_L = L;
                    *best_ms = ms;
                }
            }
            std::fclose(f);
        }
        if (*hit) {
            seg_len = *best_L;
            return build();
        }
        if ((rc = autotune(dQ, best_L, best_ms))) return rc;
        FILE* f = std::fopen(path, "a");
        if (!f) {
            set_error("cannot append to the tuning cache %s", path);
            return NEGF_EARG;
        }
        std::fprintf(f, "%s %d %.6f\n", key.c_str(), *best_L, *best_ms);
        std::fclose(f);
        return NEGF_OK;
    }

    // Read the plans' device status words into the sticky root flag
    // (synchronises the batch stream).
    int poll_status()
    {
        int s1 = 0, s2 = 0, rc;
        if ((rc = eq_plan.check_status(stream, &s1))) return rc;
        if (n_neq > 0 && (rc = neq_plan.check_status(stream, &s2))) return rc;
        root_flag |= (s1 | s2) & 1;
        return NEGF_OK;
    }
    int check_status()
    {
        int rc = poll_status();
        if (rc) return rc;
        const int f = root_flag;
        root_flag = 0;
        if (f) {
            set_error("surface Green's function: no root with Im<0 at some energy (g = 0 used)");
            return NEGF_EROOT;
        }
        return NEGF_OK;
    }

    void drop_graph()
    {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        gexec = nullptr;
        graph_ok = false;
        graph_dQ = nullptr;
    }

    template <class T>
    int upload(DevBuf& b, const std::vector<T>& v)
    {
        int rc = b.ensure(std::max<size_t>(1, v.size()) * sizeof(T));
        if (rc) return rc;
        NEGF_HIP_TRY(hipStreamSynchronize(stream));
        if (!v.empty()) NEGF_HIP_TRY(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return NEGF_OK;
    }

    std::vector<Tube*> tubes;
    std::vector<long long> off;
    std::vector<long long> versions;
    int seg_len = 0;
    int n_eq = 0, n_neq = 0;
    long long generation = 0;  // incremented by every build()
    int root_flag = 0;         // sticky: a surface GF without an Im<0 root since the last status call
    double energy_sites = 0.0;
    RgfPlan eq_plan, neq_plan;
    DevBuf eq_chains, neq_chains, d_rho0, d_eq, d_neq;
    hipStream_t stream = nullptr, stream_neq = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_in = nullptr, ev_out = nullptr, ev_fork = nullptr,
               ev_join = nullptr;
    bool timing = false, use_graph = false, graph_ok = false;
    hipGraphExec_t gexec = nullptr;
    double* graph_dQ = nullptr;
};

}  // namespace negf

using namespace negf;

struct negf_batch {
    Batch b;
};

namespace negf {
int batch_prepare(negf_batch* b, int iter, bool* rebuilt) { return b->b.prepare(iter, rebuilt); }
int batch_enqueue(negf_batch* b, double* dQ, hipStream_t st, const int* tube_active)
{
    return b->b.enqueue_on(dQ, st, tube_active);
}
long long batch_generation(const negf_batch* b) { return b->b.generation; }
int batch_poll_status(negf_batch* b) { return b->b.poll_status(); }
void batch_status_words(const negf_batch* b, const int** eq, const int** neq)
{
    *eq = b->b.eq_plan.status_ptr();
    *neq = b->b.n_neq > 0 ? b->b.neq_plan.status_ptr() : nullptr;
}
}  // namespace negf

extern "C" {

int negf_batch_create(negf_tube* const* tubes, int ntubes, negf_batch** out)
{
    clear_error();
    if (!tubes || ntubes <= 0 || !out) {
        set_error("invalid batch arguments");
        return NEGF_EARG;
    }
    std::vector<Tube*> t(ntubes);
    for (int i = 0; i < ntubes; ++i) {
        if (!tubes[i]) {
            set_error("tube %d is NULL", i);
            return NEGF_EARG;
        }
        t[i] = &tubes[i]->t;
    }
    negf_batch* h = new negf_batch;
    int rc = h->b.init(t.data(), ntubes);
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return NEGF_OK;
}

int negf_batch_destroy(negf_batch* b)
{
    if (!b) return NEGF_EARG;
    delete b;
    return NEGF_OK;
}

int negf_batch_offsets(const negf_batch* b, long long* offsets)
{
    clear_error();
    if (!b || !offsets) {
        set_error("null argument");
        return NEGF_EARG;
    }
    std::copy(b->b.off.begin(), b->b.off.end(), offsets);
    return NEGF_OK;
}

int negf_batch_solve(negf_batch* b, int iter, double* d_Q, void* stream)
{
    clear_error();
    if (!b || !d_Q) {
        set_error("null argument");
        return NEGF_EARG;
    }
    return b->b.solve(iter, d_Q, (hipStream_t)stream);
}

int negf_batch_use_graph(negf_batch* b, int on)
{
    clear_error();
    if (!b) {
        set_error("null argument");
        return NEGF_EARG;
    }
    b->b.use_graph = on != 0;
    if (!on) b->b.drop_graph();
    return NEGF_OK;
}

int negf_batch_set_segment_length(negf_batch* b, int L)
{
    clear_error();
    if (!b || L < 0) {
        set_error("invalid segment length");
        return NEGF_EARG;
    }
    b->b.seg_len = L;
    return b->b.build();
}

int negf_batch_autotune(negf_batch* b, double* d_Q, int* segment_length, double* ms)
{
    clear_error();
    if (!b || !d_Q) {
        set_error("null argument");
        return NEGF_EARG;
    }
    return b->b.autotune(d_Q, segment_length, ms);
}

int negf_batch_autotune_cached(negf_batch* b, double* d_Q, const char* cache_path, int* segment_length, double* ms,
                               int* from_cache)
{
    clear_error();
    if (!b || !d_Q || !cache_path || !segment_length || !ms || !from_cache) {
        set_error("null argument");
        return NEGF_EARG;
    }
    return b->b.autotune_cached(d_Q, cache_path, segment_length, ms, from_cache);
}

int negf_batch_enable_timing(negf_batch* b, int on)
{
    if (!b) return NEGF_EARG;
    b->b.timing = on != 0;
    return NEGF_OK;
}

int negf_batch_last_time(negf_batch* b, double* ms)
{
    clear_error();
    if (!b || !ms || !b->b.timing) {
        set_error("timing not enabled");
        return NEGF_EARG;
    }
    return b->b.last_ms(ms);
}

int negf_batch_energy_sites(const negf_batch* b, double* energy_sites)
{
    if (!b || !energy_sites) return NEGF_EARG;
    *energy_sites = b->b.energy_sites;
    return NEGF_OK;
}

int negf_batch_status(negf_batch* b)
{
    clear_error();
    if (!b) {
        set_error("null argument");
        return NEGF_EARG;
    }
    return b->b.check_status();
}

}  // extern "C"
