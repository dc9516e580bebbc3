// server.cpp -- batched and served execution of many spectrometer jobs.
//
// No counterpart in the reference (it processes one job per process).  Each
// step advances every active job by one chunk of FFTs; jobs with the same FFT
// length and window share launches (pass 1 grid z / pass 2 grid y index the
// jobs).  Each job's |X|^2 accumulator and phase-cal accumulator live on the
// device from admission to completion.  Groups alternate between two HIP
// streams so different FFT lengths overlap on the device.  A fixed job set can
// be captured once into a hipGraph (both streams, forked/joined by events) and
// replayed.
#include "../../include/swspec_hip.h"
#include "engine.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <tuple>
#include <vector>

using namespace swspec;

namespace {

constexpr int kStreams = 2;
// Per stream: the pass-1 output of one launch, which bounds the FFTs a
// two-pass job advances per step (4-channel 3.2M-point jobs: 40 at 2 GiB,
// 10 at 512 MiB; the 64-stream batch ran 3-4% faster at 2 GiB, graph and
// plain).  SWSPEC_SERVER_SCRATCH_MB overrides (64..16384).
static size_t y_scratch_bytes()
{
    static const size_t v = [] {
        const char* e = getenv("SWSPEC_SERVER_SCRATCH_MB");
        const long long mb = e ? atoll(e) : 2048;
        return (size_t)std::min(16384LL, std::max(64LL, mb)) << 20;
    }();
    return v;
}
constexpr int kMaxG = 64;                    // FFTs per job per chunk with phase-cal (mask bits)
constexpr int kMaxGPlain = 1024;             // FFTs per job per chunk without phase-cal
constexpr int kMaxGShort = 16384;            // the same for single-pass plans (no intermediate)

struct SJob {
    int64_t id = 0;
    swspec_hip_job d{};
    Plan* plan = nullptr;
    Unpack up;
    long long rawfft = 0;      // bytes per FFT
    long long nwrap = 0;       // FFTs available in the input ring
    int done = 0;              // FFTs enqueued so far
    float* acc = nullptr;      // 2*(M+1) floats
    bool pcal = false;
    PcalGeom pg;
    double* pcal_acc = nullptr;
    double* pcal_part = nullptr;
    int ntones = 0, tone_step = 0;
    int last_stream = -1;      // stream of the job's last batch (-1: none yet)
    size_t mem = 0;
    std::vector<void*> bufs;
    std::vector<size_t> buf_sizes;
};

struct Pending {
    hipEvent_t ev[kStreams];
    bool used[kStreams];
    std::vector<SJob*> finished;
    std::vector<hipEvent_t> marks;   // start-of-step marks, reusable once the step is done
};

}  // namespace

struct swspec_hip_server {
    int device = 0;
    hipStream_t st[kStreams] = {nullptr, nullptr};
    std::map<std::pair<long long, int>, Plan*> plans;
    std::map<Plan*, int> plan_seq;      // creation order: the order groups launch in each step
    std::map<std::tuple<long long, long long, int, int>, float2*> rots;
    std::vector<SJob*> active;
    std::deque<Pending> pending;
    std::vector<int64_t> done_ids;
    float2* yscr[kStreams] = {nullptr, nullptr};
    int64_t budget = 0, used = 0;
    int max_active = 256;
    int64_t next_id = 1;
    std::map<size_t, std::vector<void*>> pool;
    std::vector<hipEvent_t> free_events;
    std::vector<hipEvent_t> capture_events;   // recorded during a capture; freed after it
    std::string err;
    // counters (swspec_hip_server_stats)
    int64_t n_steps = 0, n_cross_waits = 0, n_admitted = 0, n_completed = 0;
};

static int sfail(swspec_hip_server* s, int code, const std::string& m)
{
    if (s) s->err = m;
    return code;
}

#define SCHK(s, call)                                                                             \
    do {                                                                                          \
        hipError_t _e = (call);                                                                   \
        if (_e != hipSuccess)                                                                     \
            return sfail((s), _e == hipErrorOutOfMemory ? SWSPEC_ENOMEM : SWSPEC_EDEVICE,         \
                         std::string(#call) + ": " + hipGetErrorString(_e));                     \
    } while (0)

static void* pool_get(swspec_hip_server* s, size_t bytes)
{
    auto it = s->pool.find(bytes);
    if (it != s->pool.end() && !it->second.empty()) {
        void* p = it->second.back();
        it->second.pop_back();
        return p;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    return p;
}

static void pool_put(swspec_hip_server* s, void* p, size_t bytes) { s->pool[bytes].push_back(p); }

static hipEvent_t ev_get(swspec_hip_server* s)
{
    if (!s->free_events.empty()) {
        hipEvent_t e = s->free_events.back();
        s->free_events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
    return e;
}

static void job_release(swspec_hip_server* s, SJob* j)
{
    for (size_t i = 0; i < j->bufs.size(); i++) pool_put(s, j->bufs[i], j->buf_sizes[i]);
    s->used -= (int64_t)j->mem;
    delete j;
}

extern "C" int swspec_hip_server_create(int device, int64_t budget, int max_active, swspec_hip_server** out)
{
    if (!out) return SWSPEC_EINVAL;
    *out = nullptr;
    swspec_hip_server* s = new swspec_hip_server();
    s->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete s; return SWSPEC_EDEVICE; }
    for (int i = 0; i < kStreams; i++) {
        if (hipStreamCreateWithFlags(&s->st[i], hipStreamNonBlocking) != hipSuccess) {
            swspec_hip_server_destroy(s);
            return SWSPEC_EDEVICE;
        }
        if (hipMalloc(&s->yscr[i], y_scratch_bytes()) != hipSuccess) {
            swspec_hip_server_destroy(s);
            return SWSPEC_ENOMEM;
        }
    }
    size_t fr = 0, tot = 0;
    (void)hipMemGetInfo(&fr, &tot);
    s->budget = budget > 0 ? budget : (int64_t)(fr / 2);
    s->max_active = max_active > 0 ? max_active : 256;
    *out = s;
    return SWSPEC_OK;
}

static int server_unpack(const swspec_hip_job* j, Unpack& up, long long& rawfft, std::string& why)
{
    up = Unpack();
    if (j->bits_per_sample != 2) { why = "only 2-bit samples"; return -1; }
    if (j->channel < 0 || j->channel >= j->source_channels) { why = "channel out of range"; return -1; }
    const bool raw = j->source_format == SWSPEC_FMT_RAWSIGNED || j->source_format == SWSPEC_FMT_RAWUNSIGNED;
    if (raw && j->source_channels == 1) {
        up.kind = kUnpackSingleChannel;
        rawfft = j->fft_points / 4;
        return 0;
    }
    if ((raw && (j->source_channels % 4) == 0) ||
        (j->source_format == SWSPEC_FMT_MK5B && (j->source_channels == 4 || j->source_channels == 8))) {
        const bool incr = j->source_format == SWSPEC_FMT_MK5B || j->channelorder_increasing;
        const int step = j->source_channels / 4;
        up.kind = kUnpackMultiChannel;
        up.step = step;
        up.cbyte = incr ? j->channel / 4 : (step - 1) - j->channel / 4;
        up.shift = incr ? 2 * (j->channel % 4) : 6 - 2 * (j->channel % 4);
        rawfft = j->fft_points * step;
        return 0;
    }
    why = "decoder not implemented for this format/channel count";
    return -1;
}

extern "C" int swspec_hip_server_submit(swspec_hip_server* s, const swspec_hip_job* jd, int64_t* id)
{
    if (!s || !jd) return SWSPEC_EINVAL;
    if (jd->struct_size != sizeof(swspec_hip_job)) return sfail(s, SWSPEC_EINVAL, "swspec_hip_job size mismatch");
    if (jd->nffts < 1 || !jd->d_raw || !jd->d_out_auto) return sfail(s, SWSPEC_EINVAL, "job needs nffts, input, output");
    if ((int)s->active.size() >= s->max_active) return sfail(s, SWSPEC_EAGAIN, "active job limit");
    SJob* j = new SJob();
    j->d = *jd;
    std::string why;
    if (server_unpack(jd, j->up, j->rawfft, why) != 0) { delete j; return sfail(s, SWSPEC_EUNSUPPORTED, why); }
    j->nwrap = (long long)(jd->raw_len / (uint64_t)j->rawfft);
    if (j->nwrap < 1) { delete j; return sfail(s, SWSPEC_EINVAL, "input shorter than one FFT"); }
    if (j->nwrap >= jd->nffts) j->nwrap = 0;   // no wrap needed
    // plan cache, per (length, window, column stage kind)
    const bool mfma = (jd->flags & SWSPEC_FLAG_MFMA64) != 0;
    const auto key = std::make_pair((long long)jd->fft_points, (int)jd->window | (mfma ? 0x100 : 0));
    auto it = s->plans.find(key);
    if (it == s->plans.end()) {
        Plan* p = new Plan();
        std::string e = plan_create(*p, jd->fft_points, jd->window, s->st[0], mfma);
        if (!e.empty()) { delete p; delete j; return sfail(s, SWSPEC_EUNSUPPORTED, e); }
        it = s->plans.emplace(key, p).first;
        const int seq = (int)s->plan_seq.size();
        s->plan_seq[p] = seq;
    }
    j->plan = it->second;
    const long long M = j->plan->M;
    const size_t acc_bytes = (size_t)j->plan->smax * 2 * (size_t)(M + 1) * sizeof(float);   // slices
    size_t mem = acc_bytes;
    // phase-cal geometry (swspectrometer.cpp:178-196 rules)
    if (jd->extract_pcal) {
        const double fs = jd->samplingfreq, off = jd->pcaloffsethz, harm = jd->pcalharmonicshz;
        const long long fsi = (long long)fs, offi = (long long)off;
        if (offi <= 0 || harm <= 0) { delete j; return sfail(s, SWSPEC_EINVAL, "phase-cal offset/spacing"); }
        long long a = fsi, b = offi;
        while (b) { long long t = a % b; a = b; b = t; }
        const int T = (int)(fs / harm), L = (int)(fs / (double)a);
        if (T >= 1 && T <= 1024 && L >= T && (L % T) == 0 && (jd->fft_points % L) == 0) {
            j->pcal = true;
            j->pg.T = T;
            j->pg.L = L;
            pcal_geometry(j->pg, jd->fft_points);
            const auto rk = std::make_tuple(fsi, offi, T, L);
            auto ri = s->rots.find(rk);
            if (ri == s->rots.end()) {
                std::vector<float2> rot(L);
                const double dphi = 2 * M_PI * (-off / fs);
                for (int i = 0; i < L; i++) rot[i] = make_float2((float)cos(dphi * i), (float)sin(dphi * i));
                float2* d = nullptr;
                SCHK(s, hipMalloc(&d, L * sizeof(float2)));
                SCHK(s, hipMemcpy(d, rot.data(), L * sizeof(float2), hipMemcpyHostToDevice));
                ri = s->rots.emplace(rk, d).first;
            }
            j->pg.d_rot = ri->second;
            const double bw = fs / 2;
            j->ntones = std::max(0, std::min(T, (int)std::floor((bw - off) / harm) + 1));
            j->tone_step = (int)std::floor(T * (harm / fs));
            mem += (size_t)T * 2 * sizeof(double) + (size_t)kMaxG * j->pg.nblk_max * T * 2 * sizeof(double);
        }
    }
    if (s->used + (int64_t)mem > s->budget) { delete j; return sfail(s, SWSPEC_EAGAIN, "device memory budget"); }
    // resident state
    auto alloc = [&](size_t bytes) -> void* {
        void* p = pool_get(s, bytes);
        if (p) { j->bufs.push_back(p); j->buf_sizes.push_back(bytes); }
        return p;
    };
    j->acc = (float*)alloc(acc_bytes);
    if (!j->acc) { job_release(s, j); return sfail(s, SWSPEC_ENOMEM, "accumulator allocation"); }
    SCHK(s, launch_zero(j->acc, acc_bytes, s->st[0]));
    if (j->pcal) {
        j->pcal_acc = (double*)alloc((size_t)j->pg.T * 2 * sizeof(double));
        j->pcal_part = (double*)alloc((size_t)kMaxG * j->pg.nblk_max * j->pg.T * 2 * sizeof(double));
        if (!j->pcal_acc || !j->pcal_part) { job_release(s, j); return sfail(s, SWSPEC_ENOMEM, "phase-cal allocation"); }
        SCHK(s, launch_zero(j->pcal_acc, (size_t)j->pg.T * 2 * sizeof(double), s->st[0]));
    }
    // the other stream must see the zeroed state
    hipEvent_t e = ev_get(s);
    SCHK(s, hipEventRecord(e, s->st[0]));
    SCHK(s, hipStreamWaitEvent(s->st[1], e, 0));
    s->free_events.push_back(e);
    j->mem = mem;
    s->used += (int64_t)mem;
    j->id = s->next_id++;
    s->active.push_back(j);
    s->n_admitted++;
    if (id) *id = j->id;
    return SWSPEC_OK;
}

static void retire(swspec_hip_server* s, Pending& p)
{
    for (SJob* j : p.finished) {
        s->done_ids.push_back(j->id);
        s->n_completed++;
        job_release(s, j);
    }
    for (int i = 0; i < kStreams; i++)
        if (p.used[i]) s->free_events.push_back(p.ev[i]);
    for (hipEvent_t e : p.marks) s->free_events.push_back(e);
}

// Enqueue one step.  capture: no events, no waits (stream capture).
static int step_enqueue(swspec_hip_server* s, bool capture, std::vector<SJob*>* finished_out)
{
    // group by plan, groups in plan creation order (deterministic across runs)
    std::map<int, std::pair<Plan*, std::vector<SJob*>>> groups;
    for (SJob* j : s->active) {
        auto& g = groups[s->plan_seq[j->plan]];
        g.first = j->plan;
        g.second.push_back(j);
    }
    s->n_steps++;
    int launch = 0;
    bool used[kStreams] = {false, false};
    std::vector<SJob*> finished;
    // A job's batch may land on another stream than its previous batch (the
    // batch order shifts as jobs arrive and finish).  Its accumulator and
    // phase-cal read-modify-writes must stay ordered, so each stream's tail at
    // the start of this step is marked, and a batch that takes a job from
    // stream o waits for that mark (it covers every earlier launch on o).
    hipEvent_t mark[kStreams] = {nullptr, nullptr};
    std::vector<hipEvent_t> marks;
    for (int i = 0; i < kStreams; i++) {
        mark[i] = ev_get(s);
        if (!mark[i]) return sfail(s, SWSPEC_EDEVICE, "hipEventCreate failed");
        marks.push_back(mark[i]);
        SCHK(s, hipEventRecord(mark[i], s->st[i]));
    }
    auto keep_marks = [&]() {
        if (capture) s->capture_events.insert(s->capture_events.end(), marks.begin(), marks.end());
    };
    for (auto& g : groups) {
        Plan* p = g.second.first;
        std::vector<SJob*>& js = g.second.second;
        const double per = (double)p->M * sizeof(float2);
        // jobs per launch: keep >= 8 FFTs per job per chunk when the scratch
        // allows (single-pass plans use none)
        int nj = p->single ? kMaxJobs
                           : (int)std::max(1.0, std::min((double)kMaxJobs, std::floor((double)y_scratch_bytes() / (8.0 * per))));
        for (size_t b = 0; b < js.size(); b += nj) {
            const int n = (int)std::min<size_t>(nj, js.size() - b);
            bool anyp = false;
            for (int k = 0; k < n; k++) anyp |= js[b + k]->pcal;
            const double gmax = anyp ? (double)kMaxG : (double)(p->single ? kMaxGShort : kMaxGPlain);
            const int G = p->single ? (int)gmax
                                    : (int)std::max(1.0, std::min(gmax, std::floor((double)y_scratch_bytes() / (n * per))));
            const int si = launch++ % kStreams;
            hipStream_t st = s->st[si];
            used[si] = true;
            {
                bool wait_other[kStreams] = {false, false};
                for (int k = 0; k < n; k++) {
                    SJob* j = js[b + k];
                    if (j->last_stream >= 0 && j->last_stream != si) wait_other[j->last_stream] = true;
                    j->last_stream = si;
                }
                for (int o = 0; o < kStreams; o++)
                    if (wait_other[o]) {
                        s->n_cross_waits++;
                        hipError_t we = hipStreamWaitEvent(st, mark[o], 0);
                        if (we != hipSuccess) {
                            keep_marks();
                            return sfail(s, SWSPEC_EDEVICE, std::string("hipStreamWaitEvent: ") + hipGetErrorString(we));
                        }
                    }
            }
            JobDev jd[kMaxJobs];
            int maxnf = 0;
            for (int k = 0; k < n; k++) {
                SJob* j = js[b + k];
                JobDev& J = jd[k];
                const int nf = std::min(G, j->d.nffts - j->done);
                J.raw = (const uint8_t*)j->d.d_raw;
                J.fft_stride = j->rawfft;
                J.fft0 = j->done;
                J.nwrap = j->nwrap;
                J.nf = nf;
                J.up = j->up;
                J.Y = p->single ? nullptr : s->yscr[si] + (size_t)k * G * (size_t)p->M;
                J.acc = j->acc;
                if (j->pcal) {
                    J.pcal_part = j->pcal_part;
                    J.pcal_acc = j->pcal_acc;
                    J.pcal_mask = nf >= 64 ? ~0ull : ((1ull << nf) - 1ull);
                }
                maxnf = std::max(maxnf, nf);
            }
            SCHK(s, launch_pass1(*p, jd, n, maxnf, st));
            // phase-cal: jobs with the same rotator geometry share a launch
            {
                std::vector<int> idx;
                for (int k = 0; k < n; k++) if (js[b + k]->pcal) idx.push_back(k);
                while (!idx.empty()) {
                    const PcalGeom& pg = js[b + idx[0]]->pg;
                    JobDev sub[kMaxJobs];
                    int ns = 0, mnf = 0;
                    std::vector<int> rest;
                    for (int k : idx) {
                        const PcalGeom& q = js[b + k]->pg;
                        if (q.T == pg.T && q.L == pg.L && q.d_rot == pg.d_rot) {
                            sub[ns++] = jd[k];
                            mnf = std::max(mnf, jd[k].nf);
                        } else {
                            rest.push_back(k);
                        }
                    }
                    SCHK(s, launch_pcal(pg, p->N, sub, ns, mnf, st));
                    idx.swap(rest);
                }
            }
            SCHK(s, launch_pass2(*p, jd, n, nullptr, maxnf, 1, p->smax, st));
            for (int k = 0; k < n; k++) {
                SJob* j = js[b + k];
                j->done += jd[k].nf;
                if (j->done >= j->d.nffts) {
                    SCHK(s, launch_finalize_jobs(*p, &j->acc, &j->d.d_out_auto, 1, 1.0 / (double)j->d.nffts,
                                                 p->smax, false, true, st));
                    if (j->pcal)
                        SCHK(s, launch_pcal_emit(j->pg, j->pcal_acc, j->d.d_out_pcal, j->d.d_tones, j->ntones,
                                                 j->tone_step, st));
                    finished.push_back(j);
                }
            }
        }
    }
    // remove finished jobs from the active set
    if (!finished.empty()) {
        std::vector<SJob*> keep;
        for (SJob* j : s->active)
            if (std::find(finished.begin(), finished.end(), j) == finished.end()) keep.push_back(j);
        s->active.swap(keep);
    }
    if (capture) {
        keep_marks();
        if (finished_out) finished_out->insert(finished_out->end(), finished.begin(), finished.end());
        return SWSPEC_OK;
    }
    Pending pd;
    pd.finished = finished;
    pd.marks = marks;
    for (int i = 0; i < kStreams; i++) {
        pd.used[i] = used[i];
        if (used[i]) {
            pd.ev[i] = ev_get(s);
            SCHK(s, hipEventRecord(pd.ev[i], s->st[i]));
        }
    }
    s->pending.push_back(pd);
    return SWSPEC_OK;
}

static int pending_wait_front(swspec_hip_server* s)
{
    Pending& p = s->pending.front();
    for (int i = 0; i < kStreams; i++)
        if (p.used[i]) SCHK(s, hipEventSynchronize(p.ev[i]));
    retire(s, p);
    s->pending.pop_front();
    return SWSPEC_OK;
}

extern "C" int swspec_hip_server_step(swspec_hip_server* s, int* nactive)
{
    if (!s) return SWSPEC_EINVAL;
    while (s->pending.size() >= 2) {
        int rc = pending_wait_front(s);
        if (rc) return rc;
    }
    if (!s->active.empty()) {
        int rc = step_enqueue(s, false, nullptr);
        if (rc) return rc;
    }
    if (nactive) *nactive = (int)s->active.size();
    return SWSPEC_OK;
}

extern "C" int swspec_hip_server_poll(swspec_hip_server* s, int64_t* ids, int max_ids, int* ndone)
{
    if (!s) return SWSPEC_EINVAL;
    while (!s->pending.empty()) {
        Pending& p = s->pending.front();
        bool ready = true;
        for (int i = 0; i < kStreams; i++) {
            if (!p.used[i]) continue;
            hipError_t q = hipEventQuery(p.ev[i]);
            if (q == hipErrorNotReady) { ready = false; break; }
            if (q != hipSuccess) return sfail(s, SWSPEC_EDEVICE, hipGetErrorString(q));
        }
        if (!ready) break;
        retire(s, p);
        s->pending.pop_front();
    }
    int n = 0;
    while (n < max_ids && !s->done_ids.empty()) {
        ids[n++] = s->done_ids.front();
        s->done_ids.erase(s->done_ids.begin());
    }
    if (ndone) *ndone = n;
    return SWSPEC_OK;
}

extern "C" int swspec_hip_server_sync(swspec_hip_server* s)
{
    if (!s) return SWSPEC_EINVAL;
    while (!s->pending.empty()) {
        int rc = pending_wait_front(s);
        if (rc) return rc;
    }
    for (int i = 0; i < kStreams; i++) SCHK(s, hipStreamSynchronize(s->st[i]));
    return SWSPEC_OK;
}

extern "C" int swspec_hip_server_run_batch(swspec_hip_server* s, int use_graph, int replays)
{
    if (!s) return SWSPEC_EINVAL;
    if (!use_graph) {
        while (!s->active.empty()) {
            int rc = swspec_hip_server_step(s, nullptr);
            if (rc) return rc;
        }
        return swspec_hip_server_sync(s);
    }
    int rc = swspec_hip_server_sync(s);
    if (rc) return rc;
    // capture every step of the fixed job set into one graph
    std::vector<SJob*> jobs = s->active;
    std::vector<SJob*> finished;
    hipEvent_t fork = ev_get(s), join = ev_get(s);
    SCHK(s, hipStreamBeginCapture(s->st[0], hipStreamCaptureModeGlobal));
    SCHK(s, hipEventRecord(fork, s->st[0]));
    SCHK(s, hipStreamWaitEvent(s->st[1], fork, 0));
    while (!s->active.empty()) {
        rc = step_enqueue(s, true, &finished);
        if (rc) {
            hipGraph_t g;
            (void)hipStreamEndCapture(s->st[0], &g);
            return rc;
        }
    }
    SCHK(s, hipEventRecord(join, s->st[1]));
    SCHK(s, hipStreamWaitEvent(s->st[0], join, 0));
    hipGraph_t graph = nullptr;
    SCHK(s, hipStreamEndCapture(s->st[0], &graph));
    hipGraphExec_t exec = nullptr;
    SCHK(s, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    for (int r = 0; r < std::max(1, replays); r++) SCHK(s, hipGraphLaunch(exec, s->st[0]));
    SCHK(s, hipStreamSynchronize(s->st[0]));
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    s->free_events.push_back(fork);
    s->free_events.push_back(join);
    s->free_events.insert(s->free_events.end(), s->capture_events.begin(), s->capture_events.end());
    s->capture_events.clear();
    for (SJob* j : finished) {
        s->done_ids.push_back(j->id);
        s->n_completed++;
        job_release(s, j);
    }
    return SWSPEC_OK;
}

extern "C" int swspec_hip_server_stats(const swspec_hip_server* s, int64_t* out, int n)
{
    if (!s || !out || n < 0) return SWSPEC_EINVAL;
    const int64_t v[6] = {s->n_steps, s->n_cross_waits, s->n_admitted, s->n_completed,
                          (int64_t)s->active.size(), s->used};
    for (int i = 0; i < n && i < 6; i++) out[i] = v[i];
    return SWSPEC_OK;
}

extern "C" int swspec_hip_server_destroy(swspec_hip_server* s)
{
    if (!s) return SWSPEC_EINVAL;
    for (int i = 0; i < kStreams; i++) if (s->st[i]) (void)hipStreamSynchronize(s->st[i]);
    for (SJob* j : s->active) job_release(s, j);
    s->active.clear();
    for (Pending& p : s->pending) retire(s, p);
    s->pending.clear();
    for (auto& kv : s->pool)
        for (void* p : kv.second) (void)hipFree(p);
    for (auto& kv : s->plans) { plan_destroy(*kv.second); delete kv.second; }
    for (auto& kv : s->rots) (void)hipFree(kv.second);
    for (int i = 0; i < kStreams; i++) {
        if (s->yscr[i]) (void)hipFree(s->yscr[i]);
        if (s->st[i]) (void)hipStreamDestroy(s->st[i]);
    }
    for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
    delete s;
    return SWSPEC_OK;
}

extern "C" const char* swspec_hip_server_last_error(const swspec_hip_server* s)
{
    return s ? s->err.c_str() : "";
}
