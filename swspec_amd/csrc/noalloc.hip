// noalloc.hip -- the decoder entry point for callers that cannot allocate.
//
// swspec_hip_unpack_device_noalloc is swspec_hip_unpack_device without any
// device allocation: the mark5access decodes' scratch (the VLBA modulation
// bits of a 20000-word payload and the blank-zone table of the call's payload
// segments, mark5_format_vlba.c:167-196 and blanker_mark5.c:126-200) lives in
// the file-scope array `scratch` below, one copy per GPU, held until the
// process ends.  Host code never passes the array's name to a kernel: each
// device's address comes from hipGetSymbolAddress at that device's first
// call and is kept in `g_dev`.
#include "../../include/swspec_hip.h"
#include "engine.h"

#include <hip/hip_runtime.h>

#include <mutex>

using namespace swspec;

namespace {

constexpr size_t kScratchBytes = 1u << 20;                        // per GPU
constexpr size_t kModBytes = (kVlbaModWords * 4 + 255) / 256 * 256;   // modulation bits, 256-aligned
constexpr long long kMaxSegments = (long long)((kScratchBytes - kModBytes) / sizeof(int2));

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
struct DevScratch {
    uint8_t* ptr = nullptr;   // device address of this GPU's `scratch`
    bool ready = false;       // modulation bits written
};
std::mutex g_mu;
DevScratch g_dev[kMaxDevices];

// Set-up once per device, with that device current: the array's device
// address, and the constant modulation bits copied into its head.
int scratch_for_device(int dev, uint8_t** out)
{
    if (dev < 0 || dev >= kMaxDevices) return set_library_error(SWSPEC_EDEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    DevScratch& d = g_dev[dev];
    if (!d.ready) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return set_library_error(SWSPEC_EDEVICE, "hipGetSymbolAddress(scratch) failed");
        uint32_t bits[kVlbaModWords];
        vlba_modulation_bits(bits);
        e = hipMemcpy(p, bits, sizeof(bits), hipMemcpyHostToDevice);
        if (e != hipSuccess) return set_library_error(SWSPEC_EDEVICE, "copying the VLBA modulation bits failed");
        d.ptr = (uint8_t*)p;
        d.ready = true;
    }
    *out = d.ptr;
    return SWSPEC_OK;
}

}  // namespace

extern "C" int swspec_hip_unpack_device_noalloc(const swspec_hip_cfg* cfg, const void* d_raw, int channel,
                                                int64_t count, float* d_out, void* stream)
{
    if (!cfg || !d_raw || !d_out || count < 0) return set_library_error(SWSPEC_EINVAL, "null argument");
    Unpack up;
    if (int rc = decoder_for(cfg, channel, up)) return rc;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return set_library_error(SWSPEC_EDEVICE, "hipGetDevice failed");
    const hipStream_t st = (hipStream_t)stream;
    if (is_m5a(up)) {
        // one extract_samples call over `count` samples: its blank zones see only its bytes
        const long long nseg = m5a_segments(up, count);
        if (nseg > kMaxSegments)
            return set_library_error(SWSPEC_EINVAL, "decode longer than the no-allocation scratch holds "
                                                    "(at most 130752 blank-zone entries: VLBA 1748938752 samples "
                                                    "at fanout 1 or 2, 2331918336 at fanout 4; Mark5B "
                                                    "1307520000 bytes)");
        up.vnseg = (int)nseg;
        uint8_t* base = nullptr;
        if (int rc = scratch_for_device(dev, &base)) return rc;
        up.vmod = (const uint32_t*)base;
        up.vblank = (int2*)(base + kModBytes);
        JobDev J;
        J.raw = (const uint8_t*)d_raw;
        J.raw_len = m5a_bytes(up, count);
        J.nf = 1;
        J.up = up;
        if (count > 0) {
            const hipError_t e = launch_vlba_blank(&J, 1, 1, count, st);
            if (e != hipSuccess) return set_library_error(SWSPEC_EDEVICE, hipGetErrorString(e));
        }
    }
    const hipError_t e = launch_unpack((const uint8_t*)d_raw, up, count, d_out, st);
    if (e != hipSuccess) return set_library_error(SWSPEC_EDEVICE, hipGetErrorString(e));
    return SWSPEC_OK;
}
