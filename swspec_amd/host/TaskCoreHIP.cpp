// TaskCoreHIP.cpp -- TaskCore (src/TaskCore.h:36-89) over the C ABI of
// libswspec_hip.so.  Call order and ownership follow the reference
// dispatcher (src/TaskDispatcher.cpp:92-247): all Buffers belong to main();
// the core writes outputs between run() and join(); input buffers stay
// untouched until join() returns.
#include "TaskCoreHIP.h"

#include <cstdlib>
#include <cstring>
#include <iostream>

using std::endl;

TaskCoreHIP::TaskCoreHIP(int mrank)
    : TaskCore(mrank), rank(mrank), core(NULL), cfg(NULL), buf_out(NULL), bufxpol_out(NULL), bufpcal_out(NULL),
      total_runtime(0.0), total_spectra(0), failed(false)
{
}

TaskCoreHIP::~TaskCoreHIP()
{
    if (core) swspec_hip_finalize(core);
    core = NULL;
}

void TaskCoreHIP::prepare(swspect_settings_t* settings)
{
    cfg = settings;
    std::ostream* log = settings->tlog ? static_cast<std::ostream*>(settings->tlog) : &std::cerr;

    swspec_hip_cfg c = swspec_hip_cfg();
    c.struct_size = sizeof(c);
    c.abi_version = SWSPEC_HIP_ABI_VERSION;
    c.fft_points = (int64_t)settings->fft_points;
    c.fft_overlap_factor = settings->fft_overlap_factor;
    c.fft_ssb_points = settings->fft_ssb_points;
    c.window = (int32_t)settings->wf_type;
    c.bits_per_sample = settings->bits_per_sample;
    c.source_channels = settings->source_channels;
    c.channelorder_increasing = settings->channelorder_increasing ? 1 : 0;
    c.source_format = (int32_t)settings->sourceformat;
    c.use_channel_file1 = settings->use_channel_file1;
    c.use_channel_file2 = settings->use_channel_file2;
    c.num_sources = settings->num_sources;
    c.num_xpols = settings->num_xpols;
    c.raw_fullfft_bytes = (int64_t)settings->raw_fullfft_bytes;
    c.raw_overlap_bytes = (int64_t)settings->raw_overlap_bytes;
    c.core_overlapped_ffts = settings->core_overlapped_ffts;
    c.max_spectra_per_buffer = settings->max_spectra_per_buffer;
    c.extract_pcal = settings->extract_PCal ? 1 : 0;
    c.pcal_tonebins = settings->pcal_tonebins;
    c.pcal_rotatorlen = settings->pcal_rotatorlen;
    c.samplingfreq = settings->samplingfreq;
    c.pcaloffsethz = settings->pcaloffsethz;
    c.pcalharmonicshz = settings->pcalharmonicshz;
    c.max_rawbuf_bytes = (int64_t)settings->rawbuf_size;
    std::strncpy(c.mark5_format, settings->sourceformat_str.c_str(), sizeof(c.mark5_format) - 1);
    const char* dev = getenv("SWSPEC_HIP_DEVICE");
    c.device = dev ? atoi(dev) : 0;

    if (swspec_hip_create(&c, rank, &core) != SWSPEC_OK) {
        *log << "HIP core " << rank << " prepare failed: " << swspec_hip_last_error(NULL) << endl;
        core = NULL;
        failed = true;
    }
}

int TaskCoreHIP::run(Buffer** inbuf, Buffer** outbuf, Buffer** xpolbuf, Buffer** pcalbuf)
{
    if (!core) return -1;
    const void* in[2] = {NULL, NULL};
    size_t len[2] = {0, 0};
    float* out[2] = {NULL, NULL};
    float* pc[2] = {NULL, NULL};
    float* xp[1] = {NULL};
    for (int s = 0; s < cfg->num_sources; s++) {
        in[s] = inbuf[s]->getData();
        len[s] = inbuf[s]->getLength();
        out[s] = (float*)outbuf[s]->getData();
        if (cfg->extract_PCal && pcalbuf) pc[s] = (float*)pcalbuf[s]->getData();
    }
    if (cfg->num_xpols && xpolbuf) xp[0] = (float*)xpolbuf[0]->getData();
    buf_out = outbuf;
    bufxpol_out = xpolbuf;
    bufpcal_out = pcalbuf;
    if (swspec_hip_run(core, in, len, out, xp, pc) != SWSPEC_OK) {
        std::ostream* log = cfg->tlog ? static_cast<std::ostream*>(cfg->tlog) : &std::cerr;
        *log << "HIP core " << rank << " run failed: " << swspec_hip_last_error(core) << endl;
        failed = true;
        return -1;
    }
    failed = false;
    return 0;
}

int TaskCoreHIP::join()
{
    if (!core || failed) return -1;
    int n = 0;
    size_t la = 0, lx = 0, lp = 0;
    if (swspec_hip_join(core, &n, &la, &lx, &lp) != SWSPEC_OK) return -1;
    // output lengths as TaskCoreIPP::doMaths sets them (:527-541)
    for (int s = 0; s < cfg->num_sources; s++) buf_out[s]->setLength(la);
    for (int x = 0; x < cfg->num_xpols && bufxpol_out; x++) bufxpol_out[x]->setLength(lx);
    if (cfg->extract_PCal && bufpcal_out)
        for (int s = 0; s < cfg->num_sources; s++) bufpcal_out[s]->setLength(lp);
    total_runtime += swspec_hip_last_run_seconds(core);
    total_spectra += n;
    return n;
}

int TaskCoreHIP::finalize()
{
    std::ostream* log = (cfg && cfg->tlog) ? static_cast<std::ostream*>(cfg->tlog) : &std::cerr;
    *log << "HIP core " << rank << " completed: " << total_runtime << "s total internal calculation time, "
         << total_spectra << " spectra" << endl;
    if (core) swspec_hip_finalize(core);
    core = NULL;
    return 0;
}

int TaskCoreHIP::combineResults(Buffer** outbuf, Buffer** pcalbuf)
{
    if (!core || outbuf == NULL || buf_out == NULL || pcalbuf == NULL || bufpcal_out == NULL) {
        std::cerr << "TaskCoreHIP combine : null buffer(s)!" << endl;
        return 0;
    }
    float* out[3] = {NULL, NULL, NULL};
    float* pc[2] = {NULL, NULL};
    size_t ol[3] = {0, 0, 0}, pl[2] = {0, 0};
    const int nout = cfg->num_sources + cfg->num_xpols;
    for (int i = 0; i < nout && i < 3; i++) out[i] = (float*)outbuf[i]->getData();
    if (cfg->extract_PCal)
        for (int s = 0; s < cfg->num_sources; s++) pc[s] = (float*)pcalbuf[s]->getData();
    if (swspec_hip_combine(core, out, pc, ol, pl) != SWSPEC_OK) return 0;
    for (int i = 0; i < nout && i < 3; i++) if (ol[i]) outbuf[i]->setLength(ol[i]);
    if (cfg->extract_PCal)
        for (int s = 0; s < cfg->num_sources; s++) if (pl[s]) pcalbuf[s]->setLength(pl[s]);
    return 0;
}

void TaskCoreHIP::resetBuffer(Buffer* buf)
{
    swspec_hip_reset(buf->getData(), buf->getAllocated());
}
