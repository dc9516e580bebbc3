// engine.h -- internal C++ interface of the MI355X spectrometer engine.
//
// The engine processes "jobs": one job is one channel of one raw byte stream
// (the unit that TaskCoreIPP::doMaths handles per source,
// src/IA-32/TaskCoreIPP.cpp:401-449).  Jobs that share an FFT length share one
// plan and run in the same kernel launches; cross-power pairs (:457-487) are
// job pairs processed by the same row-pass workgroups.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace swspec {

constexpr int kMaxStages = 16;
constexpr int kMaxJobs = 16;      // jobs per engine launch (kernel-argument tables)

// 2-bit decoder kinds (src/IA-32/DataUnpackerFactory.h:38-108)
enum UnpackKind : int {
    kUnpackNone = 0,
    kUnpackMultiChannel = 1,   // TwoBitUnpacker / Mk5BUnpacker: one byte group per time sample
    kUnpackSingleChannel = 2,  // TwoBitSinglechannelUnpacker: 4 samples per byte, oldest in MSB
    kUnpackVLBA = 3,           // VLBAUnpacker, one byte per time sample: VLBA1_<1|2|4>-*-4-2 / -8-1
    kUnpackMark5B = 4,         // Mark5BUnpacker: mark5access Mark5B bitstreams, decimation 1
};

// Mark5B through mark5access (Mark5BUnpacker, DataUnpackers.cpp:660-758): one
// mark5_unpack over the whole FFT, 10000-byte payloads with one blank zone
// each, no modulation; channel c of time sample t reads bit t*B + c*nbit.
constexpr int kMark5BPayload = 10000;

// VLBA framing seen by VLBAUnpacker::extract_samples (DataUnpackers.cpp:500-528)
// for the modes with one byte per time sample (ntrack = 8 x fanout, fanout
// 2^fsh): every call (one FFT) is cut into chunks of 2*(2520 - 12)*ntrack
// bytes, each an unpacker stream without headers whose payloads of
// 2500*ntrack bytes restart the modulation index; each payload is cut into
// blank zones of 32 KiB (fsh + 1 of them), and the chunk's third payload is
// its 16*ntrack-byte tail.  The modulation index counts frame words of
// `fanout` bytes, 20000 per payload.
constexpr int kVlbaPayload = 20000;    // payload bytes at 8 tracks; frame words per payload
constexpr int kVlbaChunk = 40128;      // chunk bytes at 8 tracks
constexpr int kVlbaSegPerChunk = 3;    // payloads at 0, P and the tail at 2P
constexpr int kVlbaModWords = (kVlbaPayload + 31) / 32;

struct Unpack {
    int kind = kUnpackNone;
    int step = 1;      // bytes per time sample (multi-channel)
    int cbyte = 0;     // byte offset of the channel inside a time sample
    int shift = 0;     // bit shift of the 2-bit field
    // Mk5B frames as stored on disk (multi-channel layout only): payload byte P
    // is at mk5b_frame_addr(P) of the input
    int framed = 0;
    // VLBA only (cbyte holds the channel, 0..nchan-1)
    int vfsh = 0;                        // log2 fanout: 0, 1, 2 (8, 16, 32 tracks)
    int vnbit = 2;                       // VLBA: 2 (4 channels) or 1 (8 channels); Mark5B: 1 or 2
    int vbits = 0;                       // Mark5B: bitstreams B = nchan x nbit (2..32)
    int vnseg = 0;                       // blank-zone entries per FFT (vlba_segments)
    const uint32_t* vmod = nullptr;      // modulation bits, payload index i at bit i
    int2* vblank = nullptr;              // [FFT of the chunk][segment] data zone (start, end)
};

// Mk5B framing (src/FileSource.cpp:184-215: 10016-byte frames, 16-byte header,
// stripped on the host by the reference): address of payload byte P.
constexpr long long kMk5bFrame = 10016, kMk5bHeader = 16, kMk5bPayload = 10000;
__host__ __device__ inline long long mk5b_frame_addr(long long P)
{
    const long long fr = P / kMk5bPayload;
    return fr * kMk5bFrame + kMk5bHeader + (P - fr * kMk5bPayload);
}

// Blank-zone entries of an N-sample VLBA FFT: chunks x 3 payload segments x
// (fsh + 1) zones.
inline long long vlba_segments(long long N, int fsh)
{
    const long long chunk = (long long)kVlbaChunk << fsh;
    return ((N + chunk - 1) / chunk) * kVlbaSegPerChunk * (fsh + 1);
}
// Either mark5access decoder (blank-zone tables, generic phase-cal fold).
__host__ __device__ inline bool is_m5a(const Unpack& up) { return up.kind == kUnpackVLBA || up.kind == kUnpackMark5B; }
// Blank-zone entries per FFT of N samples: VLBA as above; Mark5B one per
// 10000-byte payload of the FFT's N*B/8 bytes.
inline long long m5a_segments(const Unpack& up, long long N)
{
    if (up.kind == kUnpackMark5B) return (N * up.vbits / 8 + kMark5BPayload - 1) / kMark5BPayload;
    return vlba_segments(N, up.vfsh);
}
// Bytes of `count` samples of a mark5access decode (VLBA: one per sample).
inline long long m5a_bytes(const Unpack& up, long long count)
{
    return up.kind == kUnpackMark5B ? count * up.vbits / 8 : count;
}
// modulate[] of mark5_format_vlba.c:167-196 packed as bits.
void vlba_modulation_bits(uint32_t* words /* kVlbaModWords */);

// Stockham sub-FFT executed in LDS by one workgroup slice.
struct SubPlan {
    int L = 0;               // points
    int nst = 0;             // stages
    int radix[kMaxStages];
    int ns[kMaxStages];      // product of the radices before the stage
    unsigned mag[kMaxStages];// ceil(2^32 / ns) for q % ns by multiply-high (ns > 1)
    int twoff[kMaxStages];   // offset (in float2) of the stage's twiddles
    int T = 0;               // threads per FFT
    int twlen = 0;           // total twiddles
    // Row pass only: stage padst writes its output blocks (radix x ns
    // elements) pad elements apart beyond their length, and the last stage,
    // whose read stride is that block length, reads them with stride
    // block + pad.  Breaks up the LDS bank pattern of a stage whose ns < 16
    // and whose blocks are a multiple of 16 elements.  padst -1: none.
    int padst = -1;
    int pad = 0;
};

// Window generated in the column pass: for real sample j = 2(N2 e + c) + r,
//   theta_j = 2 pi (j + delta) / D = phi_e + psi_{c,r}
//   w_j = a0 + a1 cos(theta) + a2 sin(theta) + a3 cos(2 theta)
// phi_e comes from a per-row table (FP64-computed, FP32-stored); psi is
// computed once per thread in FP64.  The reference windows
// (TaskCoreIPP.cpp:569-605) in this form:
//   None: kind 0;  Cosine: a2 = 1, D = 2N, delta = 1/2;
//   Cosine2: a0 = 1/2, a1 = -1/2, D = N, delta = 1/2;
//   Hamming: 0.54, -0.46, D = N-1;  Hann: 1/2, -1/2, D = N-1;
//   Blackman(alpha = 0.5): 0.75, -0.5, a3 = -0.25, D = N-1.
struct WindowGen {
    int kind = 0;            // 0: rectangular
    float a0 = 1.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    double D = 1.0, delta = 0.0;
    const float2* phi = nullptr;
};

// Two-pass four-step real FFT plan for N real points (M = N/2 complex).
struct Plan {
    long long N = 0, M = 0;
    int N1 = 0, N2 = 0;      // M = N1 * N2; pass 1 = N1-point columns, pass 2 = N2-point rows
    int B = 0;               // columns per pass-1 workgroup pipeline (divides N2)
    int p1_threads = 0, p2_threads = 0;
    int p1_maxpts = 0, p2_maxpts = 0;
    int p1_cx = 0, p2_cx = 0;    // sub-plan uses composite register radices (10..32)
    int p1_fpw = 1;          // FFT pipelines per pass-1 workgroup (short columns)
    int mfma64 = 0;          // pass 1 opens with the radix-64 split-FP16 MFMA stage
    int single = 0;          // short FFTs in one pass (k_short): N1 = M, N2 = 1, sp1 the M-point FFT
    int smax = 1;            // accumulator slices per job (pass-2 FFT parallelism)
    int npairs = 0;          // pass-2 workgroups per FFT (row pairs)
    int twshift = 0;         // four-step twiddle split W_M^e = A[e>>s] * B[e & (2^s-1)]
    SubPlan sp1, sp2;
    int window = 0;
    // device tables
    float2* d_stw1 = nullptr;   // stage twiddles for sub-FFT N1
    float2* d_stw2 = nullptr;   // stage twiddles for sub-FFT N2
    float2* d_twA = nullptr;
    float2* d_twB = nullptr;
    float2* d_rA = nullptr;     // W_N^{k1}, k1 < N1
    float2* d_rB = nullptr;     // W_N^{N1 k2}, k2 < N2
    float2* d_winphi = nullptr; // N1 window row phases (cos, sin)
    float4* d_wcol = nullptr;   // window column constants (P, Q, R, S) per (column c < N2, sample r): [2c + r]
    uint4* d_mft = nullptr;     // radix-64 MFMA stage: DFT-64 A operands [8][4][hi, lo][64 lanes]
    WindowGen win;
};

// Build a plan (host tables in FP64, rounded to FP32 once).  Returns an
// error string (empty on success).
// Planner choices the autotuner varies (0 / -1: the planner's own choice).
struct PlanKnobs {
    int p1_maxpts = 0;   // column pass points per thread: 16 or 32
    int p1_bmax = 0;     // upper bound on the columns per workgroup
    int cx_mask = -1;    // composite radices: bit 0 column pass, bit 1 row pass
    int p2_maxpts = 0;   // row pass points per thread: 16 or 32
    int single = -1;     // short FFTs: 0 forces two passes (-1: the planner's choice)
};
// Measured choices per FFT length (swspec_hip_autotune); plan_geometry uses
// them when no knobs are passed.  Process-wide, thread-safe.
void plan_set_tuned(long long N, const PlanKnobs& k);
bool plan_get_tuned(long long N, PlanKnobs& k);
void plan_clear_tuned();

// mfma64: open the column FFT with a radix-64 stage on the matrix cores
// (needs a column length that is a multiple of 64).
std::string plan_create(Plan& p, long long N, int window, hipStream_t stream, bool mfma64 = false,
                        const PlanKnobs* knobs = nullptr);
void plan_destroy(Plan& p);
// Factorisation only (no device work): used by tests and the tuner.
std::string plan_geometry(Plan& p, long long N, bool mfma64 = false, const PlanKnobs* knobs = nullptr);

// One job as the kernels see it.
struct JobDev {
    const uint8_t* raw = nullptr;  // device pointer: FFT i of the stream starts at raw + i*fft_stride
    long long fft_stride = 0;      // bytes between consecutive FFT starts
    long long fft0 = 0;            // stream index of FFT 0 of the chunk
    long long nwrap = 0;           // > 0: input is a ring of nwrap FFTs (index taken mod nwrap)
    Unpack up;
    float2* Y = nullptr;           // pass-1 output [G][N1][N2]
    float* acc = nullptr;          // accumulator: (sum, comp) pairs, permuted layout, M+1 bins
    double* pcal_part = nullptr;   // pcal partial sums [G][nblk][T] double2
    double* pcal_acc = nullptr;    // pcal accumulator [T] double2
    unsigned long long pcal_mask = 0;  // bit f: FFT f of the chunk gets phase-cal
    int nf = 0;                    // FFTs of this job in the chunk (<= launch nfft)
    long long raw_len = 0;         // bytes readable from raw (VLBA blank-zone probes past an FFT)
    long long pay0 = 0;            // framed Mk5B: payload offset of FFT 0 (raw = first frame; FFT f
                                   // starts at payload byte pay0 + f*fft_stride)
};

struct XpolDev {
    float* acc = nullptr;          // (re, im, comp_re, comp_im) per bin, permuted layout
};

// Level-1 blocks per job of the phase-cal partial reduce (k_pcal_reduce1).
constexpr int kPcalRedBlocks = 32;

struct PcalGeom {
    int T = 0;        // tone bins
    int L = 0;        // rotator length
    int nblk = 0;     // blocks per FFT (generic kernel)
    int chunk = 0;    // samples per block (generic kernel)
    int threads = 0;  // generic kernel
    int fast = 0;     // 1: position-owning integer-count kernel (k_pcal_fast) usable
    int nblk_fast = 0;
    int threads_fast = 0;
    int ppb = 0;      // rotator periods per block (fast kernel)
    int gpt = 0;      // groups of 4 positions per thread (fast kernel)
    // grouped fast kernel (k_pcal_group): blocks of threads_grp threads, one
    // group of 4 positions each; npos_grp blocks cover the L/4 groups, times
    // the period chunks of ppb_grp periods (nblk_grp blocks per FFT)
    int threads_grp = 0, npos_grp = 0, ppb_grp = 0, nblk_grp = 0;
    int nblk_max = 0; // partial-sum slots per FFT to allocate
    const float2* d_rot = nullptr;  // FP32 rotator (FP64-computed), L entries
};

// Choose the phase-cal launch geometry for N-point FFTs (fills threads, nblk,
// chunk, fast, ppb, gpt from T and L).
void pcal_geometry(PcalGeom& g, long long N);

// Start of FFT f of a chunk.
__host__ __device__ inline const uint8_t* job_fft_ptr(const JobDev& J, int f)
{
    long long i = J.fft0 + f;
    if (J.nwrap > 0) i %= J.nwrap;
    return J.raw + i * J.fft_stride;
}

// Launchers (engine.hip).
hipError_t launch_pass1(const Plan& p, const JobDev* jobs, int njobs, int nfft, hipStream_t s);
// Each job's accumulator holds `smax` slices of 2*(M+1) floats (cross-power:
// 4*(M+1)); pass 2 spreads a chunk's FFTs over up to smax slices and
// finalize sums all slices in order.
hipError_t launch_pass2(const Plan& p, const JobDev* jobs, int njobs, const XpolDev* xp, int nfft,
                        int kahan, int smax, hipStream_t s);
hipError_t launch_zero(void* ptr, size_t bytes, hipStream_t s);
// dst[i] += src[i]; src[i] = 0 for i < n
hipError_t launch_add_clear(float* dst, float* src, size_t n, hipStream_t s);
hipError_t launch_pcal(const PcalGeom& g, long long N, const JobDev* jobs, int njobs, int nfft,
                       hipStream_t s);
// All jobs of a spectrum in one launch: acc[j] holds nslices slices of
// (2 floats; cross-power: 4) x (M+1) entries, out[j] the natural-order
// spectrum; zero clears what is read (no separate fill).
hipError_t launch_finalize_jobs(const Plan& p, float* const* acc, float* const* out, int njobs, double scale,
                                int nslices, bool xpol, bool zero, hipStream_t s);
hipError_t launch_pcal_emit_jobs(const PcalGeom& g, double* const* acc, float* const* out, double* const* tones,
                                 int njobs, int ntones, int tone_step, hipStream_t s);
hipError_t launch_pcal_emit(const PcalGeom& g, double* acc, float* out, double* tones, int ntones,
                            int tone_step, hipStream_t s);
hipError_t launch_unpack(const uint8_t* raw, Unpack up, long long count, float* out, hipStream_t s);
// VLBA jobs: blank zones of every payload segment of the chunk's FFTs
// (blanker_mark5.c:126-200), written to each job's up.vblank.
hipError_t launch_vlba_blank(const JobDev* jobs, int njobs, int nfft, long long N, hipStream_t s);

// Decoder for a standalone decode call of the C ABI (core.cpp): the
// factory choice for `channel` plus SWSPEC_FLAG_MK5B_FRAMED.  Returns 0 or a
// negative SWSPEC_* code with the library error text set.
int decoder_for(const struct swspec_hip_cfg* cfg, int channel, Unpack& up);
// Sets the library error text (swspec_hip_last_error(NULL)); returns code.
int set_library_error(int code, const char* msg);

}  // namespace swspec
