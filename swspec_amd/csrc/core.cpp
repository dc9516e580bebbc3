// core.cpp -- TaskCoreHIP behind the C ABI of include/swspec_hip.h.
//
// One core = one HIP stream + one FFT plan + resident accumulators.  run()
// walks a buffer with the reference's doMaths loop arithmetic
// (src/IA-32/TaskCoreIPP.cpp:394-524): FFT k of a spectrum starts
// raw_overlap_bytes after FFT k-1, a spectrum of K = core_overlapped_ffts
// FFTs is followed by a hop of (raw_fullfft_bytes - raw_overlap_bytes), a
// trailing incomplete spectrum is emitted unscaled.  The FFTs of a spectrum
// are processed in chunks of G; each chunk is pass 1 (+ phase-cal) + pass 2
// over all jobs of the core.
#include "../../include/swspec_hip.h"
#include "engine.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <strings.h>
#include <vector>

using namespace swspec;

namespace {

thread_local std::string g_lib_error;

struct SpecWork {
    long long start;      // byte offset of FFT 0 of the spectrum
    int nfft;
    bool complete;
    std::vector<unsigned char> pcal;  // [fft][job] gate (TaskCoreIPP.cpp:414-418)
};

}  // namespace

struct swspec_hip_core {
    swspec_hip_cfg cfg{};
    int rank = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // phase-cal fold on its own stream (it reads only the raw input), so it
    // runs beside the FFT passes; events order it against input and emission
    hipStream_t pstream = nullptr;
    hipEvent_t ev_in = nullptr, ev_pc = nullptr, ev_em = nullptr;
    // row pass on its own stream over two Y buffers: the column pass of chunk
    // i+1 runs beside the row pass of chunk i (one is VALU-bound, the other
    // waits on memory as much as it computes); ev_p1/ev_p2 order each buffer's
    // writer and reader
    hipStream_t s2 = nullptr;
    hipEvent_t ev_p1[2] = {nullptr, nullptr}, ev_p2[2] = {nullptr, nullptr};
    int ybufs = 1;
    Plan plan;
    PcalGeom pg;
    float2* d_rot = nullptr;
    int nsrc = 1, nx = 0, njobs = 1;
    int job_src[kMaxJobs] = {0};
    Unpack up[kMaxJobs];
    int G = 1;
    float2* d_Y = nullptr;
    float* d_acc = nullptr;       // njobs * 2*(M+1)
    float* d_xacc = nullptr;      // 4*(M+1)
    double* d_pcal_part = nullptr;
    double* d_pcal_acc = nullptr;
    int ntones = 0, tone_step = 0;
    bool vlba = false;             // some job decodes VLBA
    uint32_t* d_vmod = nullptr;    // modulation bits
    int2* d_vblank = nullptr;      // [job][FFT of the chunk][segment] blank zones
    // host-mode input copied in segments on its own stream (pinned host
    // memory makes them asynchronous): a chunk's launches wait only for the
    // segments holding its bytes, so PCIe overlaps the FFT passes
    hipStream_t cstream = nullptr;
    hipEvent_t ev_cs = nullptr;
    std::vector<hipEvent_t> ev_seg;
    size_t seg_bytes = 0;          // 0: input device-resident or copied whole
    int nseg = 0;
    int seg_waited = -1, seg_waited_p = -1;   // last segment waited for (main, phase-cal stream)
    // host-mode staging
    uint8_t* d_in[2] = {nullptr, nullptr};
    size_t d_in_cap = 0;
    float* d_out = nullptr;       // staging outputs (auto, xpol, pcal)
    size_t d_out_cap = 0;
    double* d_tones = nullptr;
    size_t d_tones_cap = 0;
    std::vector<double> h_tones;
    // last run
    float* h_auto[kMaxJobs] = {nullptr};
    float* h_xpol = nullptr;
    float* h_pcal[kMaxJobs] = {nullptr};
    float* o_auto[kMaxJobs] = {nullptr};   // where the last run wrote (device)
    float* o_xpol = nullptr;
    float* o_pcal[kMaxJobs] = {nullptr};
    bool host_mode = false;
    bool pending = false;
    int last_nspectra = 0;
    double last_seconds = 0.0;
    std::string err;
    // per-kernel profiling (HIP events around launches on `stream`)
    bool profiling = false;
    unsigned prof_mask = 0xF;        // bit k: time SWSPEC_K_ k
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<int> ev_kind;        // kernel id of event pair i
    double k_ms[4] = {0, 0, 0, 0};
    long long k_n[4] = {0, 0, 0, 0};
};

// Records an event pair around the launches issued inside its scope.
struct ProfScope {
    swspec_hip_core* c;
    hipStream_t s;
    int idx = -1;
    ProfScope(swspec_hip_core* core, int kind, hipStream_t stream = nullptr)
        : c(core), s(stream ? stream : core->stream)
    {
        if (!c->profiling || !(c->prof_mask & (1u << kind))) return;
        const size_t need = 2 * (c->ev_used + 1);
        while (c->ev_pool.size() < need) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            c->ev_pool.push_back(e);
        }
        idx = (int)c->ev_used++;
        c->ev_kind.resize(c->ev_used);
        c->ev_kind[idx] = kind;
        (void)hipEventRecord(c->ev_pool[2 * idx], s);
    }
    ~ProfScope()
    {
        if (idx >= 0) (void)hipEventRecord(c->ev_pool[2 * idx + 1], s);
    }
};

static void prof_collect(swspec_hip_core* c)
{
    for (size_t i = 0; i < c->ev_used; i++) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev_pool[2 * i], c->ev_pool[2 * i + 1]) == hipSuccess) {
            c->k_ms[c->ev_kind[i]] += ms;
            c->k_n[c->ev_kind[i]] += 1;
        }
    }
    c->ev_used = 0;
}

static int fail(swspec_hip_core* c, int code, const std::string& msg)
{
    if (c) c->err = msg; else g_lib_error = msg;
    return code;
}

static int hipfail(swspec_hip_core* c, hipError_t e, const char* what)
{
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(c, e == hipErrorOutOfMemory ? SWSPEC_ENOMEM : SWSPEC_EDEVICE, m);
}

#define HIPCHK(c, call)                                              \
    do {                                                             \
        hipError_t _e = (call);                                      \
        if (_e != hipSuccess) return hipfail((c), _e, #call);        \
    } while (0)

// ---------------------------------------------------------------------------
// settings derivation (src/swspectrometer.cpp:67-231)
// ---------------------------------------------------------------------------

static long long gcd_ll(long long a, long long b)   // src/Helpers.cpp:150-162
{
    for (;;) {
        a = a % b;
        if (a == 0) return b;
        b = b % a;
        if (b == 0) return a;
    }
}

extern "C" int swspec_hip_derive(const swspec_hip_ini* ini, swspec_hip_derived* d)
{
    if (!ini || !d) return fail(nullptr, SWSPEC_EINVAL, "null argument");
    std::memset(d, 0, sizeof(*d));
    const int cores = ini->num_cores > 0 ? ini->num_cores : 1;
    size_t maxraw;
    if (ini->max_source_buffer_mb >= 0) {
        maxraw = (size_t)ini->max_source_buffer_mb;
        maxraw *= 1024 * 1024 / 2;    // MByte, double-buffered (:104-106)
    } else {
        maxraw = (size_t)ini->platform_max_raw_mb * 1024 * 1024;
    }
    if (ini->use_channel_file1 < 1 || ini->use_channel_file1 > ini->source_channels ||
        ini->use_channel_file2 < 1 || ini->use_channel_file2 > ini->source_channels)
        return fail(nullptr, SWSPEC_EINVAL, "invalid channel number");
    if (ini->fft_points <= 0 || ini->fft_overlap_factor <= 0)
        return fail(nullptr, SWSPEC_EINVAL, "invalid FFT points or overlap");
    d->use_channel_file1 = ini->use_channel_file1 - 1;
    d->use_channel_file2 = ini->use_channel_file2 - 1;
    d->num_xpols = (ini->do_xpol && ini->num_sources == 2) ? 1 : 0;
    const float fs = (float)((double)ini->bandwidth_hz * 2.0);
    d->samplingfreq = fs;
    const size_t N = (size_t)ini->fft_points;
    {
        const float prod = fs * ini->fft_integ_seconds;   // float arithmetic (:168)
        d->averaged_ffts = (int)(prod / (float)N);
    }
    d->fft_overlap_points = (int)(N / (size_t)ini->fft_overlap_factor);
    d->fft_ssb_points = (int)(N / 2 + 1);
    const double rbpcs = (ini->bits_per_sample * ini->source_channels) / 8.0;
    d->raw_fullfft_bytes = (int)((double)N * rbpcs);
    d->raw_overlap_bytes = (int)((double)d->fft_overlap_points * rbpcs);
    d->extract_pcal = ini->extract_pcal;
    if (d->extract_pcal) {
        const float off = ini->pcal_offset_hz, harm = ini->pcal_harmonics_hz;
        if ((long long)off == 0) return fail(nullptr, SWSPEC_EINVAL, "PCalOffsetHz 0 with ExtractPCal");
        d->pcal_tonebins = (int)(fs / harm);
        d->pcal_rotatorlen = (int)(fs / (float)gcd_ll((long long)fs, (long long)off));
        d->pcal_pulses = d->pcal_tonebins ? d->pcal_rotatorlen / d->pcal_tonebins : 0;
        d->pcal_pulses_per_fft = (int)(N / (size_t)d->pcal_rotatorlen);
        if ((N % (size_t)d->pcal_rotatorlen) != 0) d->extract_pcal = 0;   // :185-189
    }
    maxraw /= (size_t)cores;
    size_t tent = (size_t)d->raw_fullfft_bytes * (size_t)d->averaged_ffts;
    if (tent > maxraw) {
        d->max_spectra_per_buffer = 0;
        d->max_specffts_per_buffer = 1;
        d->max_buffers_per_spectrum = d->averaged_ffts;
        tent = (size_t)d->raw_fullfft_bytes;
        while (tent <= maxraw && (d->max_buffers_per_spectrum % 2) == 0) {
            tent *= 2;
            d->max_buffers_per_spectrum /= 2;
            d->max_specffts_per_buffer *= 2;
        }
    } else {
        d->max_spectra_per_buffer = 1;
        d->max_specffts_per_buffer = d->averaged_ffts;
        d->max_buffers_per_spectrum = 0;
        while (2 * tent <= maxraw) {
            tent *= 2;
            d->max_spectra_per_buffer *= 2;
        }
    }
    d->rawbuf_size = (int64_t)tent;
    d->max_rawbuf_size = (int64_t)maxraw;
    d->core_averaged_ffts = d->max_specffts_per_buffer;
    d->core_overlapped_ffts =
        ini->fft_overlap_factor * d->max_specffts_per_buffer - (ini->fft_overlap_factor - 1);
    return SWSPEC_OK;
}

extern "C" int swspec_hip_cfg_from_ini(const swspec_hip_ini* ini, const swspec_hip_derived* d,
                                       swspec_hip_cfg* c)
{
    if (!ini || !d || !c) return fail(nullptr, SWSPEC_EINVAL, "null argument");
    std::memset(c, 0, sizeof(*c));
    c->struct_size = sizeof(*c);
    c->abi_version = SWSPEC_HIP_ABI_VERSION;
    c->fft_points = ini->fft_points;
    c->fft_overlap_factor = ini->fft_overlap_factor;
    c->fft_ssb_points = d->fft_ssb_points;
    c->window = ini->window;
    c->bits_per_sample = ini->bits_per_sample;
    c->source_channels = ini->source_channels;
    c->channelorder_increasing = ini->channelorder_increasing;
    c->source_format = ini->source_format;
    c->use_channel_file1 = d->use_channel_file1;
    c->use_channel_file2 = d->use_channel_file2;
    c->num_sources = ini->num_sources;
    c->num_xpols = d->num_xpols;
    c->raw_fullfft_bytes = d->raw_fullfft_bytes;
    c->raw_overlap_bytes = d->raw_overlap_bytes;
    c->core_overlapped_ffts = d->core_overlapped_ffts;
    c->max_spectra_per_buffer = d->max_spectra_per_buffer;
    c->extract_pcal = d->extract_pcal;
    c->pcal_tonebins = d->pcal_tonebins;
    c->pcal_rotatorlen = d->pcal_rotatorlen;
    c->samplingfreq = d->samplingfreq;
    c->pcaloffsethz = ini->pcal_offset_hz;
    c->pcalharmonicshz = ini->pcal_harmonics_hz;
    c->max_rawbuf_bytes = d->rawbuf_size;
    return SWSPEC_OK;
}

// ---------------------------------------------------------------------------
// decoder selection (src/IA-32/DataUnpackerFactory.h:38-108)
// ---------------------------------------------------------------------------

static int pick_unpack(const swspec_hip_cfg* c, int channel, Unpack& up, std::string& why)
{
    up = Unpack();
    if (channel < 0 || channel >= c->source_channels) { why = "channel out of range"; return -1; }
    if (c->bits_per_sample != 2 &&
        !(c->bits_per_sample == 1 && (c->source_format == SWSPEC_FMT_VLBA || c->source_format == SWSPEC_FMT_MARK5B))) {
        why = "only 2-bit samples (and 1-bit VLBA / Mark5B) are implemented";
        return -1;
    }
    const bool raw = c->source_format == SWSPEC_FMT_RAWSIGNED || c->source_format == SWSPEC_FMT_RAWUNSIGNED;
    if (raw && (c->source_channels % 4) == 0) {
        // TwoBitUnpacker (DataUnpackers.cpp:191-244)
        const int step = c->source_channels / 4;
        up.kind = kUnpackMultiChannel;
        up.step = step;
        if (c->channelorder_increasing) {
            up.cbyte = channel / 4;
            up.shift = 2 * (channel % 4);
        } else {
            up.cbyte = (step - 1) - channel / 4;
            up.shift = 6 - 2 * (channel % 4);
        }
        return 0;
    }
    if (raw && c->source_channels == 1) {
        up.kind = kUnpackSingleChannel;     // TwoBitSinglechannelUnpacker (:261-305)
        return 0;
    }
    if (c->source_format == SWSPEC_FMT_MK5B && (c->source_channels == 4 || c->source_channels == 8)) {
        // Mk5BUnpacker 4/8-channel branches (:358-407): increasing order always
        up.kind = kUnpackMultiChannel;
        up.step = c->source_channels / 4;
        up.cbyte = channel / 4;
        up.shift = 2 * (channel % 4);
        return 0;
    }
    if (c->source_format == SWSPEC_FMT_VLBA) {
        // VLBAUnpacker (:459-535) over mark5access's format string
        // "VLBA1_<fanout>-<Mbps>-<nchan>-<nbit>[/<decimation>]" (mark5_stream.c:556-571)
        char fmt[sizeof(c->mark5_format) + 1];
        std::memcpy(fmt, c->mark5_format, sizeof(c->mark5_format));
        fmt[sizeof(c->mark5_format)] = 0;
        int fan = 0, mbps = 0, nch = 0, nbit = 0, dec = 1;
        const bool prefix = strncasecmp(fmt, "VLBA1_", 6) == 0;
        const int r = prefix ? sscanf(fmt + 6, "%d-%d-%d-%d/%d", &fan, &mbps, &nch, &nbit, &dec) : 0;
        if (r < 4) { why = "VLBA format string not understood: '" + std::string(fmt) + "'"; return -1; }
        // VLBAUnpacker advances its source by the samples it decodes, not the
        // bytes (:515-522, "TODO: multiply payloadbytes by fanout"): only the
        // modes with one byte per time sample (nchan x nbit = 8) decode
        // consistently, at fanout 1, 2, 4 and without decimation
        if ((fan != 1 && fan != 2 && fan != 4) || dec != 1 || !((nch == 4 && nbit == 2) || (nch == 8 && nbit == 1))) {
            why = "VLBA decode implemented for VLBA1_<1|2|4>-<Mbps>-4-2 and VLBA1_<1|2|4>-<Mbps>-8-1 "
                  "(one byte per time sample, no decimation) only";
            return -1;
        }
        if (c->source_channels != nch || c->bits_per_sample != nbit) {
            why = "VLBA format needs SourceChannels = " + std::to_string(nch) + " and BitsPerSample = " +
                  std::to_string(nbit);
            return -1;
        }
        if (c->fft_points % fan != 0) { why = "VLBA FFT length must be a multiple of the fanout"; return -1; }
        up.kind = kUnpackVLBA;
        up.step = 1;
        up.vfsh = fan == 1 ? 0 : (fan == 2 ? 1 : 2);
        up.vnbit = nbit;
        up.cbyte = channel % nch;   // allchannels[channel % nchan] (:518)
        return 0;
    }
    if (c->source_format == SWSPEC_FMT_MARK5B) {
        // Mark5BUnpacker (:660-758) over mark5access's format string
        // "Mark5B-<Mbps>-<nchan>-<nbit>[/<decimation>]" (mark5_stream.c:556-571)
        char fmt[sizeof(c->mark5_format) + 1];
        std::memcpy(fmt, c->mark5_format, sizeof(c->mark5_format));
        fmt[sizeof(c->mark5_format)] = 0;
        int mbps = 0, nch = 0, nbit = 0, dec = 1;
        const bool prefix = strncasecmp(fmt, "Mark5B-", 7) == 0;
        const int r = prefix ? sscanf(fmt + 7, "%d-%d-%d/%d", &mbps, &nch, &nbit, &dec) : 0;
        if (r < 3) { why = "Mark5B format string not understood: '" + std::string(fmt) + "'"; return -1; }
        const int B = nch * nbit;
        // canHandleConfig (:755-758): 1 or 2 bits, even channel count; one
        // extract_samples call decodes the whole FFT, so the bytes it reads
        // (N*B/8) are the settings' raw bytes per FFT only without decimation
        if ((nbit != 1 && nbit != 2) || (nch % 2) != 0 || !(B == 2 || B == 4 || B == 8 || B == 16 || B == 32) ||
            dec != 1) {
            why = "Mark5B decode implemented for Mark5B-<Mbps>-<nchan>-<nbit> with an even channel count, "
                  "1 or 2 bits, 2..32 bitstreams and no decimation";
            return -1;
        }
        if (c->source_channels != nch || c->bits_per_sample != nbit) {
            why = "Mark5B format needs SourceChannels = " + std::to_string(nch) + " and BitsPerSample = " +
                  std::to_string(nbit);
            return -1;
        }
        if (B < 8 && (c->fft_points % (8 / B)) != 0) {
            why = "Mark5B FFT length must fill whole bytes";
            return -1;
        }
        up.kind = kUnpackMark5B;
        up.step = 1;
        up.vnbit = nbit;
        up.vbits = B;
        up.cbyte = channel % nch;   // allchannels[channel % nchan] (:734)
        return 0;
    }
    why = "decoder not implemented for this format/channel count";
    return -1;
}

// ---------------------------------------------------------------------------
// create / destroy
// ---------------------------------------------------------------------------

static int core_alloc(swspec_hip_core* c)
{
    const Plan& p = c->plan;
    const size_t M1 = (size_t)p.M + 1;
    const size_t S = (size_t)p.smax;   // accumulator slices per job
    if (!p.single)   // single-pass plans keep no intermediate
        HIPCHK(c, hipMalloc(&c->d_Y, (size_t)c->ybufs * c->njobs * c->G * (size_t)p.M * sizeof(float2)));
    HIPCHK(c, hipMalloc(&c->d_acc, (size_t)c->njobs * S * 2 * M1 * sizeof(float)));
    HIPCHK(c, (hipError_t)launch_zero(c->d_acc, (size_t)c->njobs * S * 2 * M1 * sizeof(float), c->stream));
    if (c->nx) {
        HIPCHK(c, hipMalloc(&c->d_xacc, S * 4 * M1 * sizeof(float)));
        HIPCHK(c, (hipError_t)launch_zero(c->d_xacc, S * 4 * M1 * sizeof(float), c->stream));
    }
    if (c->cfg.extract_pcal) {
        const size_t part = (size_t)c->njobs * c->G * c->pg.nblk_max * c->pg.T * 2;
        HIPCHK(c, hipMalloc(&c->d_pcal_part, part * sizeof(double)));
        HIPCHK(c, hipMalloc(&c->d_pcal_acc, (size_t)c->njobs * c->pg.T * 2 * sizeof(double)));
        HIPCHK(c, (hipError_t)launch_zero(c->d_pcal_acc, (size_t)c->njobs * c->pg.T * 2 * sizeof(double),
                                          c->stream));
    }
    if (c->vlba) {
        uint32_t bits[kVlbaModWords];
        vlba_modulation_bits(bits);
        int nseg = 1;
        for (int j = 0; j < c->njobs; j++)
            if (is_m5a(c->up[j])) nseg = (int)m5a_segments(c->up[j], c->cfg.fft_points);
        HIPCHK(c, hipMalloc(&c->d_vmod, sizeof(bits)));
        HIPCHK(c, hipMemcpy(c->d_vmod, bits, sizeof(bits), hipMemcpyHostToDevice));
        HIPCHK(c, hipMalloc(&c->d_vblank, (size_t)c->njobs * c->G * nseg * sizeof(int2)));
        for (int j = 0; j < c->njobs; j++) {
            if (!is_m5a(c->up[j])) continue;
            c->up[j].vmod = c->d_vmod;
            c->up[j].vnseg = nseg;
            c->up[j].vblank = c->d_vblank + (size_t)j * c->G * nseg;
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SWSPEC_OK;
}

static int core_setup(swspec_hip_core* c, const swspec_hip_cfg* cfg, int njobs, const int* job_src,
                      const int* job_ch)
{
    if (cfg->struct_size != sizeof(swspec_hip_cfg) || cfg->abi_version != SWSPEC_HIP_ABI_VERSION)
        return fail(c, SWSPEC_EINVAL, "swspec_hip_cfg size/version mismatch");
    c->cfg = *cfg;
    if (cfg->fft_ssb_points != cfg->fft_points / 2 + 1) return fail(c, SWSPEC_EINVAL, "fft_ssb_points mismatch");
    if (cfg->raw_overlap_bytes <= 0 || cfg->raw_fullfft_bytes < cfg->raw_overlap_bytes)
        return fail(c, SWSPEC_EINVAL, "raw byte counts invalid");
    if (cfg->core_overlapped_ffts < 1) return fail(c, SWSPEC_EINVAL, "core_overlapped_ffts < 1");
    if (cfg->window < 0 || cfg->window > 5) return fail(c, SWSPEC_EINVAL, "unknown window");
    c->njobs = njobs;
    c->nsrc = cfg->num_sources;
    c->nx = cfg->num_xpols;
    for (int j = 0; j < njobs; j++) {
        std::string why;
        c->job_src[j] = job_src[j];
        if (pick_unpack(cfg, job_ch[j], c->up[j], why) != 0) return fail(c, SWSPEC_EUNSUPPORTED, why);
        c->vlba |= is_m5a(c->up[j]);
        if (cfg->flags & SWSPEC_FLAG_MK5B_FRAMED) {
            if (cfg->source_format != SWSPEC_FMT_MK5B || c->up[j].kind != kUnpackMultiChannel)
                return fail(c, SWSPEC_EUNSUPPORTED, "SWSPEC_FLAG_MK5B_FRAMED needs Mk5B input");
            c->up[j].framed = 1;
        }
    }
    // bytes consumed per FFT must match the decoder layout
    {
        const long long per = (c->up[0].kind == kUnpackSingleChannel) ? cfg->fft_points / 4
                              : (c->up[0].kind == kUnpackMark5B)      ? cfg->fft_points * c->up[0].vbits / 8
                                                                       : cfg->fft_points * c->up[0].step;
        if (per != cfg->raw_fullfft_bytes) return fail(c, SWSPEC_EINVAL, "raw_fullfft_bytes does not match format");
        if (c->up[0].kind == kUnpackSingleChannel && (cfg->raw_overlap_bytes * 4) % 2 != 0)
            return fail(c, SWSPEC_EINVAL, "overlap not byte aligned");
    }
    if (hipSetDevice(cfg->device) != hipSuccess) return fail(c, SWSPEC_EDEVICE, "hipSetDevice failed");
    HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreate(&c->ev0));
    HIPCHK(c, hipEventCreate(&c->ev1));
    // VLBA phase-cal reads the per-chunk blank-zone table the main stream
    // rewrites, so it stays on the main stream
    // The phase-cal fold runs on the main stream between the passes: with the
    // grouped fold (each raw word read once for a source's channels) that
    // measured as fast as a second stream beside them (4.43 ms per 3.2M
    // 4-channel step either way), and the kernel keeps its own duration in a
    // trace.  SWSPEC_PCAL_SIDE=1 puts it on a stream of its own.
    static const bool pcal_side = [] {
        const char* e = getenv("SWSPEC_PCAL_SIDE");
        return e && atoi(e) != 0;
    }();
    if (cfg->extract_pcal && !c->vlba && pcal_side) {
        HIPCHK(c, hipStreamCreateWithFlags(&c->pstream, hipStreamNonBlocking));   // default priority
        for (hipEvent_t* e : {&c->ev_in, &c->ev_pc, &c->ev_em})
            HIPCHK(c, hipEventCreateWithFlags(e, hipEventDisableTiming));
    }
    std::string e = plan_create(c->plan, cfg->fft_points, cfg->window, c->stream,
                                (cfg->flags & SWSPEC_FLAG_MFMA64) != 0);
    if (!e.empty()) return fail(c, SWSPEC_EUNSUPPORTED, e);

    const long long M = c->plan.M;
    // FFTs per chunk: phase-cal masks hold 64 FFTs; without phase-cal up to
    // 1024 (single-pass plans, which keep no intermediate: 65536)
    const bool single = c->plan.single != 0;
    const int gcap = cfg->extract_pcal ? 64 : (single ? 65536 : 1024);
    if (cfg->ffts_per_chunk > 0) {
        c->G = (int)std::min<uint32_t>(cfg->ffts_per_chunk, (uint32_t)gcap);
    } else if (single) {
        c->G = gcap;
    } else {
        // chunk intermediate up to ~1 GB: the 3.2M-point four-channel case
        // (20 FFTs per chunk) measured fastest there; fewer accumulator passes
        const double per = (double)njobs * M * sizeof(float2);
        c->G = (int)std::max(1.0, std::min((double)gcap, std::floor(1.05e9 / per)));
    }
    c->G = std::max(1, std::min(c->G, cfg->core_overlapped_ffts));
    {
        // Row pass on its own stream (two Y buffers) when a spectrum spans
        // several chunks, for one-job cores and for cores with phase-cal:
        // the column pass and the phase-cal fold of the next chunk then run
        // beside the row pass.  Single-channel 3.2M cores with phase-cal:
        // +8%; the 4-channel 3.2M phase-cal step: 4.51 -> 4.32 ms (the fold
        // hidden: 4.33 with phase-cal, 4.32 without on the same streams).
        // Multi-job cores without phase-cal (4-channel 3.2M: 4.23 -> 4.32 ms)
        // and one-chunk spectra (no next chunk, only the stream handoffs)
        // keep both passes on one stream.  SWSPEC_ROW_STREAM=1/0 forces it
        // on/off.
        bool on = (c->njobs == 1 || cfg->extract_pcal) && cfg->core_overlapped_ffts > c->G;
        if (const char* e = getenv("SWSPEC_ROW_STREAM")) on = atoi(e) != 0;
        if (single) on = false;   // one pass: nothing to overlap
        if (on) {
            HIPCHK(c, hipStreamCreateWithFlags(&c->s2, hipStreamNonBlocking));   // default priority
            for (hipEvent_t* ev : {&c->ev_p1[0], &c->ev_p1[1], &c->ev_p2[0], &c->ev_p2[1]})
                HIPCHK(c, hipEventCreateWithFlags(ev, hipEventDisableTiming));
            c->ybufs = 2;
        }
    }

    if (cfg->extract_pcal) {
        const int T = cfg->pcal_tonebins, L = cfg->pcal_rotatorlen;
        if (T < 1 || L < T || (L % T) != 0 || (cfg->fft_points % L) != 0 || T > 1024)
            return fail(c, SWSPEC_EINVAL, "phase-cal geometry invalid");
        c->pg.T = T;
        c->pg.L = L;
        pcal_geometry(c->pg, cfg->fft_points);
        // the integer-count fold reads whole words of samples: VLBA (modulated,
        // blanked samples) takes the generic fold; framed input maps each
        // word's payload offset to its frame (launch_pcal checks alignment)
        if (c->vlba) c->pg.fast = 0;
        std::vector<float2> rot(L);
        const double dphi = 2 * M_PI * (-cfg->pcaloffsethz / cfg->samplingfreq);   // :130-135
        for (int i = 0; i < L; i++) {
            const double arg = dphi * (double)i;
            rot[i] = make_float2((float)cos(arg), (float)sin(arg));
        }
        HIPCHK(c, hipMalloc(&c->d_rot, L * sizeof(float2)));
        HIPCHK(c, hipMemcpy(c->d_rot, rot.data(), L * sizeof(float2), hipMemcpyHostToDevice));
        c->pg.d_rot = c->d_rot;
        const double bw = cfg->samplingfreq / 2;
        c->ntones = (int)std::floor((bw - cfg->pcaloffsethz) / cfg->pcalharmonicshz) + 1;   // PCal.cpp:328
        c->ntones = std::max(0, std::min(c->ntones, T));
        c->tone_step = (int)std::floor(T * (cfg->pcalharmonicshz / cfg->samplingfreq));     // PCal.cpp:506
    }
    return core_alloc(c);
}

extern "C" int swspec_hip_create(const swspec_hip_cfg* cfg, int rank, swspec_hip_core** out)
{
    if (!cfg || !out) return fail(nullptr, SWSPEC_EINVAL, "null argument");
    *out = nullptr;
    if (cfg->num_sources < 1 || cfg->num_sources > 2) return fail(nullptr, SWSPEC_EINVAL, "num_sources must be 1 or 2");
    if (cfg->num_xpols && cfg->num_sources != 2) return fail(nullptr, SWSPEC_EINVAL, "cross-power needs 2 sources");
    swspec_hip_core* c = new swspec_hip_core();
    c->rank = rank;
    int src[2] = {0, 1};
    int ch[2] = {cfg->use_channel_file1, cfg->use_channel_file2};
    int rc = core_setup(c, cfg, cfg->num_sources, src, ch);
    if (rc != SWSPEC_OK) {
        g_lib_error = c->err;
        swspec_hip_finalize(c);
        return rc;
    }
    *out = c;
    return SWSPEC_OK;
}

// Several channels of source 0 in one pass (no TaskCore counterpart: the
// reference runs one core per UseFile1Channel value over the same file).
extern "C" int swspec_hip_create_multichannel(const swspec_hip_cfg* cfg, int nch, const int* channels, int rank,
                                              swspec_hip_core** out)
{
    if (!cfg || !out || !channels) return fail(nullptr, SWSPEC_EINVAL, "null argument");
    *out = nullptr;
    if (nch < 1 || nch > kMaxJobs) return fail(nullptr, SWSPEC_EINVAL, "1..8 channels");
    if (cfg->num_xpols) return fail(nullptr, SWSPEC_EINVAL, "no cross-power in multichannel mode");
    swspec_hip_cfg c1 = *cfg;
    c1.num_sources = 1;
    swspec_hip_core* c = new swspec_hip_core();
    c->rank = rank;
    int src[kMaxJobs] = {0};
    int rc = core_setup(c, &c1, nch, src, channels);
    if (rc != SWSPEC_OK) {
        g_lib_error = c->err;
        swspec_hip_finalize(c);
        return rc;
    }
    *out = c;
    return SWSPEC_OK;
}

extern "C" int swspec_hip_finalize(swspec_hip_core* c)
{
    if (!c) return fail(nullptr, SWSPEC_EINVAL, "null core");
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->s2) (void)hipStreamSynchronize(c->s2);
    if (c->pstream) (void)hipStreamSynchronize(c->pstream);
    if (c->cstream) (void)hipStreamSynchronize(c->cstream);
    plan_destroy(c->plan);
    for (void* p : {(void*)c->d_rot, (void*)c->d_Y, (void*)c->d_acc, (void*)c->d_xacc, (void*)c->d_pcal_part,
                    (void*)c->d_pcal_acc, (void*)c->d_in[0], (void*)c->d_in[1], (void*)c->d_out, (void*)c->d_tones,
                    (void*)c->d_vmod, (void*)c->d_vblank})
        if (p) (void)hipFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->pstream) (void)hipStreamSynchronize(c->pstream);
    if (c->s2) (void)hipStreamSynchronize(c->s2);
    for (hipEvent_t e : {c->ev_in, c->ev_pc, c->ev_em, c->ev_p1[0], c->ev_p1[1], c->ev_p2[0], c->ev_p2[1]})
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->ev_seg) (void)hipEventDestroy(e);
    if (c->ev_cs) (void)hipEventDestroy(c->ev_cs);
    if (c->cstream) (void)hipStreamDestroy(c->cstream);
    if (c->pstream) (void)hipStreamDestroy(c->pstream);
    if (c->s2) (void)hipStreamDestroy(c->s2);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SWSPEC_OK;
}

// ---------------------------------------------------------------------------
// run
// ---------------------------------------------------------------------------

// The doMaths loop arithmetic (TaskCoreIPP.cpp:394-524).
static std::vector<SpecWork> schedule(const swspec_hip_core* c, size_t len)
{
    const swspec_hip_cfg& g = c->cfg;
    std::vector<SpecWork> out;
    size_t remain = len;
    long long off = 0;
    int acc = 0, curr = 0;
    SpecWork cur{0, 0, false, {}};
    const size_t full = (size_t)g.raw_fullfft_bytes, ovl = (size_t)g.raw_overlap_bytes;
    while (remain >= full) {
        if (cur.nfft == 0) cur.start = off;
        for (int j = 0; j < c->njobs; j++) {
            // curr_ffts counts once per source FFT (:415, :435); multichannel
            // jobs of one source share the source's count
            const int srcidx = c->job_src[j];
            const int cnt = curr + srcidx;
            cur.pcal.push_back(g.extract_pcal && (cnt % g.fft_overlap_factor) == 0);
        }
        curr += c->nsrc;
        cur.nfft++;
        off += (long long)ovl;
        remain -= ovl;
        if (++acc >= g.core_overlapped_ffts) {
            cur.complete = true;
            out.push_back(cur);
            cur = SpecWork{0, 0, false, {}};
            acc = 0;
            off += (long long)(full - ovl);
            remain -= (full - ovl);
        }
    }
    if (acc > 0) out.push_back(cur);   // partial, unscaled (:519-524)
    return out;
}

// Bytes the doMaths loop walks: framed input counts the payload bytes of its
// whole frames (headers and a trailing partial frame excluded).
static size_t payload_len(const swspec_hip_core* c, size_t len)
{
    if (c->cfg.flags & SWSPEC_FLAG_MK5B_FRAMED) return (len / (size_t)kMk5bFrame) * (size_t)kMk5bPayload;
    return len;
}

static int run_common(swspec_hip_core* c, const uint8_t* const* d_in, size_t len, float* const* d_auto,
                      float* d_xpol, float* const* d_pcal, double* d_tones)
{
    const Plan& p = c->plan;
    const swspec_hip_cfg& g = c->cfg;
    const bool framed = (g.flags & SWSPEC_FLAG_MK5B_FRAMED) != 0;
    const size_t plen = payload_len(c, len);
    std::vector<SpecWork> sched = schedule(c, plen);
    const size_t M1 = (size_t)p.M + 1;
    const int ssb = g.fft_ssb_points;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    hipStream_t pst = c->pstream ? c->pstream : c->stream;   // phase-cal stream
    if (c->pstream) {
        // input copies (host mode) are queued on the main stream before this
        HIPCHK(c, hipEventRecord(c->ev_in, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->pstream, c->ev_in, 0));
    }
    hipStream_t rst = c->s2 ? c->s2 : c->stream;   // row-pass stream
    // segmented host input: stream st waits for the segment holding byte need-1
    auto wait_input = [&](hipStream_t st, int& waited, size_t need) -> hipError_t {
        if (!c->seg_bytes || c->nseg < 1) return hipSuccess;
        const int k = need == 0 ? 0 : (int)std::min<size_t>((need - 1) / c->seg_bytes, (size_t)c->nseg - 1);
        if (k <= waited) return hipSuccess;
        waited = k;
        return hipStreamWaitEvent(st, c->ev_seg[k], 0);
    };
    long long ci = 0;                             // chunk counter: Y buffer ci % ybufs
    int ylast = 0;
    for (size_t sp = 0; sp < sched.size(); sp++) {
        const SpecWork& w = sched[sp];
        for (int f0 = 0; f0 < w.nfft; f0 += c->G, ci++) {
            const int nf = std::min(c->G, w.nfft - f0);
            const int yb = (int)(ci % c->ybufs);
            ylast = yb;
            // the row pass of chunk ci-2 has finished reading this Y buffer
            if (c->s2 && ci >= 2) HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_p2[yb], 0));
            JobDev jobs[kMaxJobs];
            bool any_pcal = false;
            for (int j = 0; j < c->njobs; j++) {
                JobDev& J = jobs[j];
                const long long off = w.start + (long long)f0 * g.raw_overlap_bytes;
                if (framed) {
                    J.raw = d_in[c->job_src[j]];
                    J.pay0 = off;
                    J.raw_len = (long long)len;
                } else {
                    J.raw = d_in[c->job_src[j]] + off;
                    J.raw_len = (long long)len - off;
                }
                J.fft_stride = g.raw_overlap_bytes;
                J.nf = nf;
                J.up = c->up[j];
                J.Y = p.single ? nullptr : c->d_Y + ((size_t)yb * c->njobs + j) * c->G * (size_t)p.M;
                J.acc = c->d_acc + (size_t)j * p.smax * 2 * M1;
                if (g.extract_pcal) {
                    J.pcal_part = c->d_pcal_part + (size_t)j * c->G * c->pg.nblk_max * c->pg.T * 2;
                    J.pcal_acc = c->d_pcal_acc + (size_t)j * c->pg.T * 2;
                    unsigned long long m = 0;
                    for (int f = 0; f < nf; f++)
                        if (w.pcal[(size_t)(f0 + f) * c->njobs + j]) m |= 1ull << f;
                    J.pcal_mask = m;
                    any_pcal |= (m != 0);
                }
            }
            // bytes the chunk reads (plus a word-alignment margin): the
            // segmented copy is used only for plain payload layouts
            const size_t need = std::min(len, (size_t)(w.start + (long long)(f0 + nf - 1) * g.raw_overlap_bytes +
                                                       g.raw_fullfft_bytes + 64));
            HIPCHK(c, wait_input(c->stream, c->seg_waited, need));
            {
                ProfScope ps(c, SWSPEC_K_PASS1);
                if (c->vlba) HIPCHK(c, launch_vlba_blank(jobs, c->njobs, nf, g.fft_points, c->stream));
                HIPCHK(c, launch_pass1(p, jobs, c->njobs, nf, c->stream));
            }
            if (c->s2) {   // rows after their columns
                HIPCHK(c, hipEventRecord(c->ev_p1[yb], c->stream));
                HIPCHK(c, hipStreamWaitEvent(c->s2, c->ev_p1[yb], 0));
            }
            if (any_pcal) {
                HIPCHK(c, wait_input(pst, pst == c->stream ? c->seg_waited : c->seg_waited_p, need));
                ProfScope ps(c, SWSPEC_K_PCAL, pst);
                HIPCHK(c, launch_pcal(c->pg, g.fft_points, jobs, c->njobs, nf, pst));
            }
            XpolDev xp;
            xp.acc = c->d_xacc;
            {
                ProfScope ps(c, SWSPEC_K_PASS2, rst);
                HIPCHK(c, launch_pass2(p, jobs, c->njobs, c->nx ? &xp : nullptr, nf,
                                       (g.flags & SWSPEC_FLAG_NO_KAHAN) ? 0 : 1, p.smax, rst));
            }
            if (c->s2) HIPCHK(c, hipEventRecord(c->ev_p2[yb], c->s2));
        }
        // the spectrum's rows are summed before finalize reads and clears the
        // accumulators (later rows wait for later columns, queued after it)
        if (c->s2) HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_p2[ylast], 0));
        ProfScope psf(c, SWSPEC_K_FINALIZE);
        // TaskCoreIPP.cpp:491-498 scale 1/K for complete spectra; partial unscaled
        const double scale = w.complete ? 1.0 / (double)g.core_overlapped_ffts : 1.0;
        {   // every job in one launch; the accumulators read are cleared in it
            float* accs[kMaxJobs];
            float* outs[kMaxJobs];
            for (int j = 0; j < c->njobs; j++) {
                accs[j] = c->d_acc + (size_t)j * p.smax * 2 * M1;
                outs[j] = d_auto[j] + sp * (size_t)ssb;
            }
            HIPCHK(c, launch_finalize_jobs(p, accs, outs, c->njobs, scale, p.smax, false, true, c->stream));
        }
        if (c->nx) {
            float* xout = d_xpol + sp * 2 * (size_t)ssb;
            HIPCHK(c, launch_finalize_jobs(p, &c->d_xacc, &xout, 1, scale, p.smax, true, true, c->stream));
        }
        if (g.extract_pcal) {
            if (c->pstream) {   // the spectrum's folds are done before they are emitted
                HIPCHK(c, hipEventRecord(c->ev_pc, c->pstream));
                HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_pc, 0));
            }
            double* paccs[kMaxJobs];
            float* pouts[kMaxJobs];
            double* touts[kMaxJobs];
            for (int j = 0; j < c->njobs; j++) {
                paccs[j] = c->d_pcal_acc + (size_t)j * c->pg.T * 2;
                pouts[j] = d_pcal && d_pcal[j] ? d_pcal[j] + sp * 2 * (size_t)c->pg.T : nullptr;
                touts[j] = d_tones ? d_tones + ((size_t)j * sched.size() + sp) * 2 * c->ntones : nullptr;
            }
            HIPCHK(c, launch_pcal_emit_jobs(c->pg, paccs, pouts, touts, c->njobs, c->ntones, c->tone_step,
                                            c->stream));
            if (c->pstream) {   // the next spectrum's folds add to the cleared accumulators
                HIPCHK(c, hipEventRecord(c->ev_em, c->stream));
                HIPCHK(c, hipStreamWaitEvent(c->pstream, c->ev_em, 0));
            }
        }
    }
    if (c->pstream) {   // all folds done before the run completes (join, input reuse)
        HIPCHK(c, hipEventRecord(c->ev_pc, c->pstream));
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_pc, 0));
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->last_nspectra = (int)sched.size();
    return SWSPEC_OK;
}

static int ensure_tones(swspec_hip_core* c, size_t nspectra)
{
    if (!c->cfg.extract_pcal || c->ntones == 0) return SWSPEC_OK;
    const size_t need = (size_t)c->njobs * std::max<size_t>(nspectra, 1) * 2 * c->ntones;
    if (need > c->d_tones_cap) {
        if (c->d_tones) HIPCHK(c, hipFree(c->d_tones));
        c->d_tones = nullptr;
        HIPCHK(c, hipMalloc(&c->d_tones, need * sizeof(double)));
        c->d_tones_cap = need;
    }
    return SWSPEC_OK;
}

static size_t min_len(const swspec_hip_core* c, const size_t* in_len)
{
    size_t len = in_len[0];
    for (int s = 1; s < c->nsrc; s++) len = std::min(len, in_len[s]);
    return len;
}

extern "C" int swspec_hip_run_device(swspec_hip_core* c, const void* const* d_in, const size_t* in_len,
                                     float* const* d_out_auto, float* const* d_out_xpol, float* const* d_out_pcal)
{
    if (!c || !d_in || !in_len || !d_out_auto) return fail(c, SWSPEC_EINVAL, "null argument");
    if (c->pending) return fail(c, SWSPEC_ESTATE, "run while a run is pending");
    if (c->nx && (!d_out_xpol || !d_out_xpol[0])) return fail(c, SWSPEC_EINVAL, "cross-power output missing");
    const size_t len = min_len(c, in_len);
    const size_t nsp = schedule(c, payload_len(c, len)).size();
    int rc = ensure_tones(c, nsp);
    if (rc) return rc;
    const uint8_t* ins[2] = {(const uint8_t*)d_in[0], c->nsrc > 1 ? (const uint8_t*)d_in[1] : nullptr};
    float* autos[kMaxJobs];
    float* pcals[kMaxJobs];
    for (int j = 0; j < c->njobs; j++) {
        autos[j] = d_out_auto[j];
        pcals[j] = d_out_pcal ? d_out_pcal[j] : nullptr;
    }
    c->host_mode = false;
    rc = run_common(c, ins, len, autos, c->nx ? d_out_xpol[0] : nullptr, pcals, c->d_tones);
    if (rc) return rc;
    for (int j = 0; j < c->njobs; j++) {   // what swspec_hip_combine_device adds up
        c->o_auto[j] = autos[j];
        c->o_pcal[j] = pcals[j];
    }
    c->o_xpol = c->nx ? d_out_xpol[0] : nullptr;
    c->pending = true;
    return SWSPEC_OK;
}

extern "C" int swspec_hip_run(swspec_hip_core* c, const void* const* in, const size_t* in_len,
                              float* const* out_auto, float* const* out_xpol, float* const* out_pcal)
{
    if (!c || !in || !in_len || !out_auto) return fail(c, SWSPEC_EINVAL, "null argument");
    if (c->pending) return fail(c, SWSPEC_ESTATE, "run while a run is pending");
    if (c->nx && (!out_xpol || !out_xpol[0])) return fail(c, SWSPEC_EINVAL, "cross-power output missing");
    const size_t len = min_len(c, in_len);
    const size_t nsp = schedule(c, payload_len(c, len)).size();   // what run_common emits
    const size_t nalloc = std::max<size_t>(nsp, 1);
    const size_t ssb = (size_t)c->cfg.fft_ssb_points;
    const size_t T = c->cfg.extract_pcal ? (size_t)c->pg.T : 0;
    // staging
    size_t inmax = 0;
    for (int s = 0; s < c->nsrc; s++) inmax = std::max(inmax, in_len[s]);
    if (inmax > c->d_in_cap) {
        for (int s = 0; s < 2; s++) if (c->d_in[s]) { HIPCHK(c, hipFree(c->d_in[s])); c->d_in[s] = nullptr; }
        for (int s = 0; s < c->nsrc; s++) HIPCHK(c, hipMalloc(&c->d_in[s], inmax));
        c->d_in_cap = inmax;
    }
    const size_t per_auto = nalloc * ssb, per_x = 2 * nalloc * ssb, per_p = 2 * nalloc * T;
    const size_t need = c->njobs * per_auto + (c->nx ? per_x : 0) + c->njobs * per_p;
    if (need > c->d_out_cap) {
        if (c->d_out) HIPCHK(c, hipFree(c->d_out));
        c->d_out = nullptr;
        HIPCHK(c, hipMalloc(&c->d_out, need * sizeof(float)));
        c->d_out_cap = need;
    }
    int rc = ensure_tones(c, nsp);
    if (rc) return rc;
    // Input copies: segments of SWSPEC_H2D_SEG_MB (default 32 MB; 0 = one
    // copy on the main stream) on the copy stream, one event per segment.
    // VLBA (the blanker probes past an FFT) and framed input (frame-addressed
    // reads) wait for the whole input.
    size_t seg = (size_t)32 << 20;
    if (const char* e = getenv("SWSPEC_H2D_SEG_MB")) seg = (size_t)std::max(0L, atol(e)) << 20;
    c->seg_bytes = 0;
    c->nseg = 0;
    c->seg_waited = c->seg_waited_p = -1;
    if (seg == 0 || c->vlba || (c->cfg.flags & SWSPEC_FLAG_MK5B_FRAMED) || inmax == 0) {
        for (int s = 0; s < c->nsrc; s++)
            HIPCHK(c, hipMemcpyAsync(c->d_in[s], in[s], in_len[s], hipMemcpyHostToDevice, c->stream));
    } else {
        if (!c->cstream) {
            HIPCHK(c, hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking));   // default priority
            HIPCHK(c, hipEventCreateWithFlags(&c->ev_cs, hipEventDisableTiming));
        }
        const int nseg = (int)((inmax + seg - 1) / seg);
        while ((int)c->ev_seg.size() < nseg) {
            hipEvent_t e = nullptr;
            HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
            c->ev_seg.push_back(e);
        }
        // the staging buffers are free once the main stream's earlier work is
        HIPCHK(c, hipEventRecord(c->ev_cs, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->cstream, c->ev_cs, 0));
        for (int k = 0; k < nseg; k++) {
            const size_t lo = (size_t)k * seg;
            for (int s = 0; s < c->nsrc; s++)
                if (lo < in_len[s])
                    HIPCHK(c, hipMemcpyAsync(c->d_in[s] + lo, (const uint8_t*)in[s] + lo,
                                             std::min(seg, in_len[s] - lo), hipMemcpyHostToDevice, c->cstream));
            HIPCHK(c, hipEventRecord(c->ev_seg[k], c->cstream));
        }
        c->seg_bytes = seg;
        c->nseg = nseg;
    }
    float* autos[kMaxJobs];
    float* pcals[kMaxJobs];
    float* base = c->d_out;
    for (int j = 0; j < c->njobs; j++) { autos[j] = base; base += per_auto; }
    float* xp = nullptr;
    if (c->nx) { xp = base; base += per_x; }
    for (int j = 0; j < c->njobs; j++) { pcals[j] = T ? base : nullptr; base += per_p; }
    const uint8_t* ins[2] = {c->d_in[0], c->d_in[1]};
    rc = run_common(c, ins, len, autos, xp, pcals, c->d_tones);
    if (rc) return rc;
    // every input copy is done before join can return (the caller may then
    // refill its buffer)
    if (c->seg_bytes && c->nseg > 0 && c->seg_waited < c->nseg - 1) {
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_seg[c->nseg - 1], 0));
        c->seg_waited = c->nseg - 1;
    }
    c->seg_bytes = 0;
    // results back to the caller's host arrays
    for (int j = 0; j < c->njobs; j++) {
        if (nsp) HIPCHK(c, hipMemcpyAsync(out_auto[j], autos[j], nsp * ssb * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        c->h_auto[j] = out_auto[j];
        c->o_auto[j] = autos[j];
        if (T && out_pcal && out_pcal[j]) {
            if (nsp) HIPCHK(c, hipMemcpyAsync(out_pcal[j], pcals[j], nsp * 2 * T * sizeof(float),
                                              hipMemcpyDeviceToHost, c->stream));
            c->h_pcal[j] = out_pcal[j];
        } else {
            c->h_pcal[j] = nullptr;
        }
    }
    if (c->nx) {
        if (nsp) HIPCHK(c, hipMemcpyAsync(out_xpol[0], xp, nsp * 2 * ssb * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        c->h_xpol = out_xpol[0];
    }
    c->host_mode = true;
    c->pending = true;
    return SWSPEC_OK;
}

extern "C" int swspec_hip_join(swspec_hip_core* c, int* nspectra, size_t* la, size_t* lx, size_t* lp)
{
    if (!c) return fail(nullptr, SWSPEC_EINVAL, "null core");
    if (!c->pending) return fail(c, SWSPEC_ESTATE, "join without run");
    c->pending = false;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) c->last_seconds = ms * 1e-3;
    if (c->profiling) prof_collect(c);
    const int n = c->last_nspectra;
    if (c->cfg.extract_pcal && c->ntones > 0 && n > 0) {
        c->h_tones.resize((size_t)c->njobs * n * 2 * c->ntones);
        HIPCHK(c, hipMemcpy(c->h_tones.data(), c->d_tones, c->h_tones.size() * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        c->h_tones.clear();
    }
    if (nspectra) *nspectra = n;
    if (la) *la = (size_t)n * c->cfg.fft_ssb_points * sizeof(float);
    if (lx) *lx = c->nx ? (size_t)n * 2 * c->cfg.fft_ssb_points * sizeof(float) : 0;
    if (lp) *lp = c->cfg.extract_pcal ? (size_t)n * c->cfg.pcal_tonebins * 2 * sizeof(float) : 0;
    return SWSPEC_OK;
}

extern "C" int swspec_hip_combine(swspec_hip_core* c, float* const* out, float* const* pcal, size_t* out_len,
                                  size_t* pcal_len)
{
    // TaskCoreIPP::combineResults (:311-349) on the host arrays of the last run
    if (!c || !out) return fail(c, SWSPEC_EINVAL, "null argument");
    if (!c->host_mode) return fail(c, SWSPEC_ESTATE, "combine needs a host-mode run");
    const size_t ssb = (size_t)c->cfg.fft_ssb_points;
    for (int s = 0; s < c->njobs; s++) {
        if (!out[s] || !c->h_auto[s]) continue;
        for (size_t k = 0; k < ssb; k++) out[s][k] += c->h_auto[s][k];
        if (out_len) out_len[s] = sizeof(float) * ssb;
        std::memset(c->h_auto[s], 0, sizeof(float) * ssb);
    }
    for (int x = 0; x < c->nx; x++) {
        float* dst = out[c->njobs + x];
        if (!dst || !c->h_xpol) continue;
        for (size_t k = 0; k < 2 * ssb; k++) dst[k] += c->h_xpol[k];
        if (out_len) out_len[c->njobs + x] = sizeof(float) * ssb;   // Q5: float-sized length (:333)
        std::memset(c->h_xpol, 0, sizeof(float) * 2 * ssb);
    }
    if (c->cfg.extract_pcal && pcal) {
        const size_t T = (size_t)c->cfg.pcal_tonebins;
        for (int s = 0; s < c->njobs; s++) {
            if (!pcal[s] || !c->h_pcal[s]) continue;
            for (size_t k = 0; k < 2 * T; k++) pcal[s][k] += c->h_pcal[s][k];
            if (pcal_len) pcal_len[s] = sizeof(float) * 2 * T;
            std::memset(c->h_pcal[s], 0, sizeof(float) * 2 * T);
        }
    }
    return SWSPEC_OK;
}

extern "C" void swspec_hip_reset(void* buf, size_t bytes)
{
    if (buf) std::memset(buf, 0, (bytes / sizeof(float)) * sizeof(float));
}

extern "C" int swspec_hip_combine_device(swspec_hip_core* c, float* const* d_out, float* const* d_pcal,
                                         size_t* out_len, size_t* pcal_len)
{
    // TaskCoreIPP::combineResults (:311-349) on device arrays: the first
    // spectrum of the last device-mode run is added into the caller's
    // combined arrays and cleared, on the core's stream; returns once done
    if (!c || !d_out) return fail(c, SWSPEC_EINVAL, "null argument");
    if (c->pending) return fail(c, SWSPEC_ESTATE, "combine while a run is pending");
    if (c->host_mode) return fail(c, SWSPEC_ESTATE, "combine_device needs a device-mode run");
    const size_t ssb = (size_t)c->cfg.fft_ssb_points;
    for (int s = 0; s < c->njobs; s++) {
        if (!d_out[s] || !c->o_auto[s]) continue;
        HIPCHK(c, launch_add_clear(d_out[s], c->o_auto[s], ssb, c->stream));
        if (out_len) out_len[s] = sizeof(float) * ssb;
    }
    for (int x = 0; x < c->nx; x++) {
        float* dst = d_out[c->njobs + x];
        if (!dst || !c->o_xpol) continue;
        HIPCHK(c, launch_add_clear(dst, c->o_xpol, 2 * ssb, c->stream));
        if (out_len) out_len[c->njobs + x] = sizeof(float) * ssb;   // Q5: float-sized length (:333)
    }
    if (c->cfg.extract_pcal && d_pcal) {
        const size_t T = (size_t)c->cfg.pcal_tonebins;
        for (int s = 0; s < c->njobs; s++) {
            if (!d_pcal[s] || !c->o_pcal[s]) continue;
            HIPCHK(c, launch_add_clear(d_pcal[s], c->o_pcal[s], 2 * T, c->stream));
            if (pcal_len) pcal_len[s] = sizeof(float) * 2 * T;
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return SWSPEC_OK;
}

extern "C" int swspec_hip_reset_device(void* d_buf, size_t bytes)
{
    // TaskCoreIPP::resetBuffer (:297-300) on device memory, by a fill kernel
    if (!d_buf) return fail(nullptr, SWSPEC_EINVAL, "null buffer");
    hipError_t e = launch_zero(d_buf, (bytes / sizeof(float)) * sizeof(float), nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return hipfail(nullptr, e, "swspec_hip_reset_device");
    return SWSPEC_OK;
}

extern "C" const char* swspec_hip_last_error(const swspec_hip_core* c)
{
    return c ? c->err.c_str() : g_lib_error.c_str();
}

extern "C" int swspec_hip_pcal_tones(swspec_hip_core* c, int source, int spectrum, double* tones, int ntones)
{
    if (!c || !tones) return fail(c, SWSPEC_EINVAL, "null argument");
    if (!c->cfg.extract_pcal) return fail(c, SWSPEC_ESTATE, "phase-cal not enabled");
    if (source < 0 || source >= c->njobs || spectrum < 0 || spectrum >= c->last_nspectra)
        return fail(c, SWSPEC_EINVAL, "source/spectrum out of range");
    if (ntones != c->ntones) return fail(c, SWSPEC_EINVAL, "ntones mismatch");
    const double* src = c->h_tones.data() + ((size_t)source * c->last_nspectra + spectrum) * 2 * c->ntones;
    std::memcpy(tones, src, sizeof(double) * 2 * c->ntones);
    return SWSPEC_OK;
}

extern "C" int swspec_hip_pcal_ntones(const swspec_hip_core* c) { return c ? c->ntones : -1; }

extern "C" double swspec_hip_last_run_seconds(const swspec_hip_core* c) { return c ? c->last_seconds : 0.0; }

extern "C" int swspec_hip_set_profiling(swspec_hip_core* c, int on)
{
    if (!c) return fail(nullptr, SWSPEC_EINVAL, "null core");
    if (c->pending) return fail(c, SWSPEC_ESTATE, "set_profiling while a run is pending");
    if (on != 0 && on != 1 && !(on & SWSPEC_PROF_MASK)) return fail(c, SWSPEC_EINVAL, "set_profiling: on");
    c->profiling = on != 0;
    c->prof_mask = (on & SWSPEC_PROF_MASK) ? (unsigned)on & 0xFu : 0xFu;
    c->ev_used = 0;
    for (int k = 0; k < 4; k++) { c->k_ms[k] = 0.0; c->k_n[k] = 0; }
    return SWSPEC_OK;
}

extern "C" int swspec_hip_kernel_stats(const swspec_hip_core* c, int kernel, double* total_ms, int64_t* launches)
{
    if (!c || kernel < 0 || kernel > 3) return fail(nullptr, SWSPEC_EINVAL, "bad argument");
    if (total_ms) *total_ms = c->k_ms[kernel];
    if (launches) *launches = c->k_n[kernel];
    return SWSPEC_OK;
}

extern "C" void* swspec_hip_host_alloc(size_t bytes)
{
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

extern "C" void swspec_hip_host_free(void* p)
{
    if (p) (void)hipHostFree(p);
}

namespace swspec {
// Decoder of `channel` for a standalone decode call (swspec_hip_unpack_device
// and its no-allocation variant): the factory choice plus the framed flag.
// Returns 0, or a negative code with the library error text set.
int decoder_for(const swspec_hip_cfg* cfg, int channel, Unpack& up)
{
    std::string why;
    if (pick_unpack(cfg, channel, up, why) != 0) return fail(nullptr, SWSPEC_EUNSUPPORTED, why);
    if (cfg->flags & SWSPEC_FLAG_MK5B_FRAMED) {
        if (cfg->source_format != SWSPEC_FMT_MK5B || up.kind != kUnpackMultiChannel)
            return fail(nullptr, SWSPEC_EUNSUPPORTED, "SWSPEC_FLAG_MK5B_FRAMED needs Mk5B input");
        up.framed = 1;   // count samples from payload byte 0 of the frames at d_raw
    }
    return SWSPEC_OK;
}

int set_library_error(int code, const char* msg) { return fail(nullptr, code, msg); }
}  // namespace swspec

extern "C" int swspec_hip_unpack_device(const swspec_hip_cfg* cfg, const void* d_raw, int channel, int64_t count,
                                        float* d_out)
{
    if (!cfg || !d_raw || !d_out || count < 0) return fail(nullptr, SWSPEC_EINVAL, "null argument");
    Unpack up;
    if (int rc = decoder_for(cfg, channel, up)) return rc;
    hipError_t e = hipSuccess;
    uint32_t* d_mod = nullptr;
    int2* d_blank = nullptr;
    if (is_m5a(up)) {
        // one extract_samples call over `count` samples: its blank zones see only its bytes
        uint32_t bits[kVlbaModWords];
        vlba_modulation_bits(bits);
        up.vnseg = (int)m5a_segments(up, count);
        e = hipMalloc(&d_mod, sizeof(bits));
        if (e == hipSuccess) e = hipMemcpy(d_mod, bits, sizeof(bits), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc(&d_blank, (size_t)std::max(up.vnseg, 1) * sizeof(int2));
        up.vmod = d_mod;
        up.vblank = d_blank;
        JobDev J;
        J.raw = (const uint8_t*)d_raw;
        J.raw_len = m5a_bytes(up, count);
        J.nf = 1;
        J.up = up;
        if (e == hipSuccess && count > 0) e = launch_vlba_blank(&J, 1, 1, count, nullptr);
    }
    if (e == hipSuccess) e = launch_unpack((const uint8_t*)d_raw, up, count, d_out, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (d_mod) (void)hipFree(d_mod);
    if (d_blank) (void)hipFree(d_blank);
    if (e != hipSuccess) return fail(nullptr, SWSPEC_EDEVICE, hipGetErrorString(e));
    return SWSPEC_OK;
}

extern "C" int swspec_hip_plan_info(int64_t fft_points, uint32_t flags, int32_t* info, int ninfo)
{
    Plan p;
    std::string e = plan_geometry(p, fft_points, (flags & SWSPEC_FLAG_MFMA64) != 0);
    if (!e.empty()) return fail(nullptr, SWSPEC_EUNSUPPORTED, e);
    int32_t v[16] = {p.N1,        p.N2,        p.B,      p.p1_threads, p.sp2.T * 2, p.npairs,    p.p1_maxpts,
                     p.p2_maxpts, p.p1_cx,     p.p2_cx,  p.sp1.nst,    p.sp2.nst,   p.mfma64,    p.sp1.radix[0],
                     p.single,    p.sp2.padst >= 0 ? p.sp2.pad : 0};
    for (int i = 0; i < ninfo && i < 16; i++) info[i] = v[i];
    return SWSPEC_OK;
}

// ---------------------------------------------------------------------------
// autotuner: measured plan choice per FFT length
// ---------------------------------------------------------------------------

// Every distinct plan the planner's knobs give for fft_points (column / row
// points per thread 16 or 32, composite radices per pass, a cap on the
// columns per workgroup) runs pass 1 + pass 2 over nfft FFTs of seeded
// single-channel 2-bit input on `device`; a candidate whose spectrum differs
// from the planner's own plan by more than 1e-5 (normwise) is rejected.  The
// fastest (minimum of 3 timed repetitions) is recorded for plan_create.
extern "C" int swspec_hip_autotune(int device, int64_t fft_points, int32_t window, int32_t nfft, int32_t* knobs_out,
                                   double* ms_best, double* ms_default, int32_t* ncand)
{
    if (fft_points < 16 || nfft < 1 || nfft > 1024 || window < 0 || window > 5)
        return fail(nullptr, SWSPEC_EINVAL, "autotune arguments");
    if (hipSetDevice(device) != hipSuccess) return fail(nullptr, SWSPEC_EDEVICE, "hipSetDevice failed");
    std::vector<PlanKnobs> cands(1);   // [0]: the planner's own choice (explicit: no tuned entry)
    for (int p1 : {16, 32})
        for (int cx = 0; cx < 4; cx++)
            for (int bm : {0, 8, 4, 2})
                for (int p2 : {16, 32}) {
                    PlanKnobs k;
                    k.p1_maxpts = p1;
                    k.cx_mask = cx;
                    k.p1_bmax = bm;
                    k.p2_maxpts = p2;
                    cands.push_back(k);
                }
    auto sig = [](const Plan& p) {
        std::string s = std::to_string(p.N1) + "/" + std::to_string(p.B) + "/" + std::to_string(p.p1_threads) + "/" +
                        std::to_string(p.p1_maxpts) + "/" + std::to_string(p.p2_maxpts) + "/" +
                        std::to_string(p.p2_threads) + ":";
        for (int i = 0; i < p.sp1.nst; i++) s += std::to_string(p.sp1.radix[i]) + ",";
        s += ":";
        for (int i = 0; i < p.sp2.nst; i++) s += std::to_string(p.sp2.radix[i]) + ",";
        return s;
    };
    std::vector<PlanKnobs> uniq;
    std::vector<std::string> seen;
    for (size_t i = 0; i < cands.size(); i++) {
        Plan g;
        if (!plan_geometry(g, fft_points, false, &cands[i]).empty()) {
            if (i == 0) return fail(nullptr, SWSPEC_EUNSUPPORTED, "no plan for this FFT length");
            continue;
        }
        const std::string sg = sig(g);
        if (std::find(seen.begin(), seen.end(), sg) != seen.end()) continue;
        seen.push_back(sg);
        uniq.push_back(cands[i]);
    }
    const long long N = fft_points, M = N / 2;
    const size_t rawfft = (size_t)N / 4;   // 4 two-bit samples per byte
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    uint8_t* d_raw = nullptr;
    float2* d_Y = nullptr;
    float* d_acc = nullptr;
    float* d_out = nullptr;
    hipError_t he = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    if (he == hipSuccess) he = hipMalloc(&d_raw, rawfft * (size_t)nfft);
    if (he == hipSuccess) he = hipMalloc(&d_Y, (size_t)nfft * (size_t)M * sizeof(float2));
    if (he == hipSuccess) he = hipMalloc(&d_out, (size_t)(M + 1) * sizeof(float));
    if (he == hipSuccess) {   // seeded bytes (xorshift), copied once
        std::vector<uint8_t> h(rawfft * (size_t)nfft);
        uint64_t x = 0x9E3779B97F4A7C15ull ^ (uint64_t)N;
        for (auto& b : h) { x ^= x << 13; x ^= x >> 7; x ^= x << 17; b = (uint8_t)(x >> 32); }
        he = hipMemcpy(d_raw, h.data(), h.size(), hipMemcpyHostToDevice);
    }
    std::vector<float> ref, got((size_t)M + 1);
    double best = 1e300, tdef = 0.0;
    int ibest = -1, tried = 0;
    size_t acc_cap = 0;
    for (size_t i = 0; he == hipSuccess && i < uniq.size(); i++) {
        Plan p;
        if (!plan_create(p, N, window, st, false, &uniq[i]).empty()) { plan_destroy(p); continue; }
        const size_t accb = (size_t)p.smax * 2 * (size_t)(M + 1) * sizeof(float);
        if (accb > acc_cap) {
            if (d_acc) (void)hipFree(d_acc);
            d_acc = nullptr;
            he = hipMalloc(&d_acc, accb);
            if (he != hipSuccess) { plan_destroy(p); break; }
            acc_cap = accb;
        }
        JobDev J;
        J.raw = d_raw;
        J.fft_stride = (long long)rawfft;
        J.raw_len = (long long)(rawfft * (size_t)nfft);
        J.up.kind = kUnpackSingleChannel;
        J.Y = d_Y;
        J.acc = d_acc;
        J.nf = nfft;
        bool ok = launch_zero(d_acc, accb, st) == hipSuccess;
        float* accs[1] = {d_acc};
        float* outs[1] = {d_out};
        // one checked run, then 3 timed repetitions (accumulators cleared by finalize)
        ok = ok && launch_pass1(p, &J, 1, nfft, st) == hipSuccess &&
             launch_pass2(p, &J, 1, nullptr, nfft, 1, p.smax, st) == hipSuccess &&
             launch_finalize_jobs(p, accs, outs, 1, 1.0, p.smax, false, true, st) == hipSuccess &&
             hipMemcpyAsync(got.data(), d_out, got.size() * sizeof(float), hipMemcpyDeviceToHost, st) == hipSuccess &&
             hipStreamSynchronize(st) == hipSuccess;
        if (ok && i == 0) ref = got;
        if (ok && i > 0) {
            double num = 0.0, den = 0.0;
            for (size_t k = 0; k < got.size(); k++) {
                const double d = (double)got[k] - (double)ref[k];
                num += d * d;
                den += (double)ref[k] * ref[k];
            }
            ok = den > 0.0 && std::sqrt(num / den) <= 1e-5;
        }
        double tmin = 1e300;
        for (int r = 0; ok && r < 3; r++) {
            float ms = 0.f;
            ok = hipEventRecord(e0, st) == hipSuccess && launch_pass1(p, &J, 1, nfft, st) == hipSuccess &&
                 launch_pass2(p, &J, 1, nullptr, nfft, 1, p.smax, st) == hipSuccess &&
                 hipEventRecord(e1, st) == hipSuccess &&
                 launch_finalize_jobs(p, accs, outs, 1, 1.0, p.smax, false, true, st) == hipSuccess &&
                 hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
            if (ok) tmin = std::min(tmin, (double)ms);
        }
        (void)hipStreamSynchronize(st);
        plan_destroy(p);
        if (!ok) {
            if (i == 0) { he = hipErrorUnknown; break; }
            (void)hipGetLastError();
            continue;
        }
        tried++;
        if (i == 0) tdef = tmin;
        if (tmin < best) { best = tmin; ibest = (int)i; }
    }
    for (void* q : {(void*)d_raw, (void*)d_Y, (void*)d_acc, (void*)d_out})
        if (q) (void)hipFree(q);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (he != hipSuccess || ibest < 0) return fail(nullptr, SWSPEC_EDEVICE, "autotune: the planner's own plan failed");
    const PlanKnobs k = uniq[(size_t)ibest];
    plan_set_tuned(N, k);
    if (knobs_out) {
        knobs_out[0] = k.p1_maxpts;
        knobs_out[1] = k.p1_bmax;
        knobs_out[2] = k.cx_mask;
        knobs_out[3] = k.p2_maxpts;
    }
    if (ms_best) *ms_best = best;
    if (ms_default) *ms_default = tdef;
    if (ncand) *ncand = tried;
    return SWSPEC_OK;
}

extern "C" void swspec_hip_autotune_clear(void) { plan_clear_tuned(); }
