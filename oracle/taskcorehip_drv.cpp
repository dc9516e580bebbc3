// Driver that runs the C++ TaskCoreHIP adapter (swspec_amd/host/TaskCoreHIP.cpp,
// the class a maintainer selects through the dispatcher's Platform_TaskCore
// switch) on the GPU the way the reference's main() and TaskDispatcher drive
// a core.  Built by `make -C oracle adapter` into oracle/_ref/ together with
// the reference's own Buffer (src/Buffer.cpp) and FileSink (src/FileSink.cpp),
// read where they lie; it is test infrastructure: a GPU test compares the
// files it writes with those of the Python dispatcher (swspec_amd/dispatcher.py).
//
//   taskcorehip_drv <out prefix> <input 1> <input 2 | -> [Key=value ...]
//
// Keys are the INI [Spectrometer] keys the core reads (numeric enums for
// WindowType and SourceFormat); the settings are derived with
// swspec_hip_derive.  The loop restates src/TaskDispatcher.cpp:36-247 for one
// core: buffer allocation as src/swspectrometer.cpp:365-413, raw reads as
// FileSource::read for unframed formats (src/FileSource.cpp:224-250: a short
// read sets end of file), combineResults when a spectrum spans several
// buffers, otherwise each buffer's spectra written directly.  Outputs:
// <prefix>_s<i>.bin, <prefix>_x0.bin, <prefix>_p<i>.bin (Binary sinks).
#include "Buffer.h"
#include "FileSink.h"
#include "Settings.h"
#include "TaskCoreHIP.h"
#include "swspec_hip.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

namespace {

struct RawSource {
    std::ifstream f;
    bool eof = false;
    bool open(const char* p)
    {
        f.open(p, std::ios::in | std::ios::binary);
        return f.is_open();
    }
    // FileSource::read, unframed formats: fill the allocation, length = bytes read
    void read(Buffer* b)
    {
        f.read(b->getData(), (std::streamsize)b->getAllocated());
        if (!f || !f.good()) eof = true;
        b->setLength((size_t)f.gcount());
    }
};

double num(const std::map<std::string, std::string>& k, const char* key, double dflt)
{
    auto it = k.find(key);
    return it == k.end() ? dflt : std::atof(it->second.c_str());
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s out_prefix in1 in2|- [Key=value ...]\n", argv[0]);
        return 2;
    }
    const std::string prefix = argv[1];
    const bool two = std::strcmp(argv[3], "-") != 0;
    std::map<std::string, std::string> k;
    for (int a = 4; a < argc; a++) {
        const char* eq = std::strchr(argv[a], '=');
        if (eq) k[std::string(argv[a], eq - argv[a])] = eq + 1;
    }
    // INI keys -> derived settings (swspectrometer.cpp:67-231 restated by the library)
    swspec_hip_ini ini = swspec_hip_ini();
    ini.num_cores = 1;
    ini.max_source_buffer_mb = num(k, "MaxSourceBufferMB", -1.0);
    ini.platform_max_raw_mb = 32;
    ini.fft_points = (int64_t)num(k, "FFTpoints", 320000);
    ini.fft_integ_seconds = (float)num(k, "FFTIntegrationTimeSec", 20.0);
    ini.fft_overlap_factor = (int)num(k, "FFToverlapFactor", 2);
    ini.window = (int)num(k, "WindowType", SWSPEC_WIN_COSINE2);
    ini.bandwidth_hz = (float)num(k, "BandwidthHz", 8e6);
    ini.pcal_offset_hz = (float)num(k, "PCalOffsetHz", 10e3);
    ini.pcal_harmonics_hz = 1e6f;
    ini.source_format = (int)num(k, "SourceFormat", SWSPEC_FMT_RAWSIGNED);
    ini.bits_per_sample = (int)num(k, "BitsPerSample", 2);
    ini.channelorder_increasing = (int)num(k, "ChannelOrderIncreasing", 1);
    ini.source_channels = (int)num(k, "SourceChannels", 1);
    ini.use_channel_file1 = (int)num(k, "UseFile1Channel", 1);
    ini.use_channel_file2 = (int)num(k, "UseFile2Channel", 1);
    ini.extract_pcal = (int)num(k, "ExtractPCal", 0);
    ini.num_sources = two ? 2 : 1;
    ini.do_xpol = (int)num(k, "DoCrossPolarization", 0);
    swspec_hip_derived d;
    if (swspec_hip_derive(&ini, &d) != SWSPEC_OK) {
        std::fprintf(stderr, "derive: %s\n", swspec_hip_last_error(NULL));
        return 3;
    }

    swspect_settings_t set;
    set.num_cores = 1;
    set.fft_points = (size_t)ini.fft_points;
    set.fft_integ_seconds = ini.fft_integ_seconds;
    set.fft_overlap_factor = ini.fft_overlap_factor;
    set.wf_type = (WindowFunctionType)ini.window;
    set.bits_per_sample = ini.bits_per_sample;
    set.channelorder_increasing = ini.channelorder_increasing != 0;
    set.source_channels = ini.source_channels;
    set.use_channel_file1 = d.use_channel_file1;
    set.use_channel_file2 = d.use_channel_file2;
    set.samplingfreq = (float)d.samplingfreq;
    set.pcaloffsethz = ini.pcal_offset_hz;
    set.pcalharmonicshz = ini.pcal_harmonics_hz;
    set.extract_PCal = d.extract_pcal != 0;
    set.sourceformat = (InputFormat)ini.source_format;
    auto fs = k.find("SourceFormatStr");
    set.sourceformat_str = fs == k.end() ? std::string() : fs->second;
    set.num_sources = ini.num_sources;
    set.num_xpols = d.num_xpols;
    set.num_sinks = set.num_sources + set.num_xpols;
    set.fft_ssb_points = d.fft_ssb_points;
    set.raw_fullfft_bytes = (size_t)d.raw_fullfft_bytes;
    set.raw_overlap_bytes = (size_t)d.raw_overlap_bytes;
    set.fft_bytes_ssb = sizeof(float) * (size_t)d.fft_ssb_points;
    set.fft_bytes_xpol = 2 * set.fft_bytes_ssb;
    set.rawbuf_size = (size_t)d.rawbuf_size;
    set.max_rawbuf_size = (size_t)d.max_rawbuf_size;
    set.max_spectra_per_buffer = d.max_spectra_per_buffer;
    set.max_specffts_per_buffer = d.max_specffts_per_buffer;
    set.max_buffers_per_spectrum = d.max_buffers_per_spectrum;
    set.core_overlapped_ffts = d.core_overlapped_ffts;
    set.core_averaged_ffts = d.core_averaged_ffts;
    set.pcal_tonebins = d.pcal_tonebins;
    set.pcal_rotatorlen = d.pcal_rotatorlen;
    set.pcal_result_bytes = set.extract_PCal ? 2 * sizeof(float) * (size_t)d.pcal_tonebins : 0;
    set.tlog = NULL;   // TaskCoreHIP logs to std::cerr then

    RawSource src[2];
    if (!src[0].open(argv[2]) || (two && !src[1].open(argv[3]))) {
        std::fprintf(stderr, "cannot open input\n");
        return 4;
    }
    const int nsrc = set.num_sources, nx = set.num_xpols;
    // main(): raw[2][sources], out[sources], xpol[xpols], pcal[sources] (one core)
    Buffer* raw[2][2];
    for (int b = 0; b < 2; b++)
        for (int s = 0; s < nsrc; s++) raw[b][s] = new Buffer(set.rawbuf_size);
    const size_t nsp = (size_t)(set.max_spectra_per_buffer > 1 ? set.max_spectra_per_buffer : 1);
    Buffer* out[2];
    Buffer* xpol[1];
    Buffer* pcal[2];
    for (int s = 0; s < nsrc; s++) {
        out[s] = new Buffer(nsp * set.fft_bytes_ssb);
        pcal[s] = new Buffer(nsp * set.pcal_result_bytes);
    }
    for (int x = 0; x < nx; x++) xpol[x] = new Buffer(nsp * set.fft_bytes_xpol);

    // sinks: auto per source, cross, phase-cal per source
    std::vector<FileSink*> sinks, psinks;
    set.sinkformat = Binary;
    for (int s = 0; s < nsrc; s++) {
        sinks.push_back(new FileSink(&set));
        sinks.back()->open(prefix + "_s" + std::to_string(s) + ".bin");
    }
    for (int x = 0; x < nx; x++) {
        sinks.push_back(new FileSink(&set));
        sinks.back()->open(prefix + "_x" + std::to_string(x) + ".bin");
    }
    if (set.extract_PCal)
        for (int s = 0; s < nsrc; s++) {
            psinks.push_back(new FileSink(&set));
            psinks.back()->open(prefix + "_p" + std::to_string(s) + ".bin");
        }

    // TaskDispatcher::TaskDispatcher: prefill, core, combined buffers
    for (int s = 0; s < nsrc; s++) src[s].read(raw[0][s]);
    TaskCoreHIP* core = new TaskCoreHIP(0);
    core->prepare(&set);
    const size_t peak = set.fft_bytes_xpol > set.fft_bytes_ssb ? set.fft_bytes_xpol : set.fft_bytes_ssb;
    std::vector<Buffer*> comb, combp;
    for (int i = 0; i < set.num_sinks; i++) {
        comb.push_back(new Buffer(peak));
        comb.back()->setLength(peak);
        core->resetBuffer(comb.back());
    }
    for (int s = 0; s < nsrc; s++) {
        combp.push_back(new Buffer(set.pcal_result_bytes));
        combp.back()->setLength(set.pcal_result_bytes);
        core->resetBuffer(combp.back());
    }

    // TaskDispatcher::run
    int dbl = 0, num_combined = 0, total = 0, rc = 0;
    bool got_eof = false;
    do {
        core->run(raw[dbl], out, xpol, pcal);
        for (int s = 0; s < nsrc; s++) src[s].read(raw[(dbl + 1) % 2][s]);
        const int n = core->join();
        if (n < 0) {
            std::fprintf(stderr, "join failed\n");
            rc = 5;
            break;
        }
        if (set.max_buffers_per_spectrum > 1) {
            core->combineResults(comb.data(), combp.data());
            if (++num_combined == set.max_buffers_per_spectrum) {
                for (int i = 0; i < set.num_sinks; i++) {
                    sinks[i]->write(comb[i]);
                    core->resetBuffer(comb[i]);
                }
                if (set.extract_PCal)
                    for (int s = 0; s < nsrc; s++) {
                        psinks[s]->write(combp[s]);
                        core->resetBuffer(combp[s]);
                    }
                total++;
                num_combined = 0;
            }
        } else {
            for (int s = 0; s < nsrc; s++) sinks[s]->write(out[s]);
            for (int x = 0; x < nx; x++) sinks[nsrc + x]->write(xpol[x]);
            if (set.extract_PCal)
                for (int s = 0; s < nsrc; s++) {
                    psinks[s]->write(pcal[s]);
                    core->resetBuffer(pcal[s]);
                }
            total += set.max_spectra_per_buffer;
        }
        dbl = (dbl + 1) % 2;
        for (int s = 0; s < nsrc; s++) {
            if (src[s].eof) got_eof = true;
            if (raw[dbl][s]->getLength() > 0) break;
        }
    } while (!got_eof);
    core->finalize();
    for (FileSink* f : sinks) f->close();
    for (FileSink* f : psinks) f->close();
    std::printf("spectra %d\n", total);
    delete core;
    return rc;
}
