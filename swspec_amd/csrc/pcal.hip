// pcal.hip -- the phase-cal extractor classes of src/IA-32/PhaseCal/PCal.cpp
// (PCal::getNew, PCalExtractorTrivial / Shifting / ImplicitShift) on the
// device, behind the swspec_hip_pcal_* entry points of include/swspec_hip.h.
//
// Every class reduces to one sample fold with 1 add per sample:
//   * Trivial / ImplicitShift add real samples into nbins bins, sample n of a
//     call going to bin (pcal_index + n) mod nbins (PCal.cpp:270-291,
//     :609-637; the reference's fixed-destination blocks into a 2*nbins
//     buffer plus the fold at getFinalPCal come to this).
//   * Shifting multiplies sample n by rotator[rotator_index + (n mod L)]
//     (L = rotatorlen, an FP32 table of 2L entries, :353-359) and adds it to
//     bin (pcal_index + (n mod L) mod nbins).  In whole rotator periods only
//     floor(L/nbins) pulses of nbins samples are added (:432-441): positions
//     n mod L >= floor(L/nbins)*nbins of those periods are dropped; the
//     trailing partial period is added whole (:444-460).  Since the rotator
//     value depends on n mod L only, the samples are first summed per
//     position n mod L and each position sum is rotated once.
// The fold (k_pfold) is a streaming read of aligned float4s: a workgroup
// owns 256 consecutive slots of the period and a group of periods, each
// lane sums its 4 slots over its share of the group in FP64, and the
// workgroup writes one partial per slot; k_preduce sums the partials per slot in a fixed
// order, k_pfinish maps slots to bins (rotating for Shifting) into the FP64
// accumulator.  getFinalPCal's DFT is
// evaluated only at the tone bins it copies (k_ptones, FP64, 1/nbins).
// Results are deterministic: no atomics, fixed summation orders.
#include "../../include/swspec_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace {

thread_local std::string g_pcal_error;

constexpr int PF_T = 256;               // threads per fold workgroup: 4 waves on the same 256 slots
constexpr int PF_TARGET_WG = 2048;      // fold workgroups aimed for (8 per CU)
constexpr int PF_MIN_MP = 4096;         // slots per period at least (short periods are repeated)
constexpr int PF_PMIN = 32;             // periods per group at least (bounds the partial traffic)
constexpr size_t PF_PART_MAX = size_t(1) << 25;   // partial doubles kept at most (256 MB)

// The fold runs over the aligned index m = n + a, where a = (x mod 16 bytes)
// / 4, so that a lane's 4 consecutive slots u0..u0+3 (u = m mod MP, MP a
// multiple of 4) come from one aligned float4 in every period.  The 4 waves
// of a workgroup take every 4th period of the group and are summed through
// LDS.  part[y * MP + u] = sum of the samples 0 <= n < len with
// (n + a) mod MP == u in the periods of group y, leaving out n < end whose
// slot is dropped (Shifting: (u - a) mod L >= jmax; jmax == L keeps all).
__global__ __launch_bounds__(PF_T) void k_pfold(const float* __restrict__ x, int a, long long len, int MP, int L,
                                                int jmax, long long end, int PG, long long P,
                                                double* __restrict__ part)
{
    __shared__ double sh[4][4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int u0 = (blockIdx.x * 64 + lane) * 4;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (u0 < MP) {
        bool drop[4];
#pragma unroll
        for (int k = 0; k < 4; k++) drop[k] = jmax < L && ((u0 + k - a) % L + L) % L >= jmax;
        const bool anydrop = drop[0] || drop[1] || drop[2] || drop[3];
        const long long p0 = (long long)blockIdx.y * PG;
        const long long p1 = min(P, p0 + PG);
#pragma unroll 4
        for (long long p = p0 + w; p < p1; p += 4) {
            const long long n0 = p * MP + u0 - a;   // x + n0 is 16-byte aligned
            float v[4];
            if (n0 >= 0 && n0 + 3 < len) {
                const float4 q = *reinterpret_cast<const float4*>(x + n0);
                v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
            } else {
#pragma unroll
                for (int k = 0; k < 4; k++) v[k] = (n0 + k >= 0 && n0 + k < len) ? x[n0 + k] : 0.f;
            }
            if (anydrop) {
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (drop[k] && n0 + k < end) v[k] = 0.f;
            }
#pragma unroll
            for (int k = 0; k < 4; k++) acc[k] += (double)v[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) sh[w][k][lane] = acc[k];
    __syncthreads();
    if (w == 0 && u0 < MP) {
#pragma unroll
        for (int k = 0; k < 4; k++)
            part[(size_t)blockIdx.y * MP + u0 + k] = ((sh[0][k][lane] + sh[1][k][lane]) + sh[2][k][lane]) + sh[3][k][lane];
    }
}

// pos[c] = sum_y part[y * MP + c]: 16 y-strided partial sums per column
// (16 waves of a workgroup share 64 columns), combined in a fixed order.
__global__ __launch_bounds__(1024) void k_preduce(const double* __restrict__ part, int Y, int MP,
                                                  double* __restrict__ pos)
{
    __shared__ double sh[16][64];
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int col = blockIdx.x * 64 + c;
    double s = 0.0;
    if (col < MP) {
        int y = g;
        for (; y + 48 < Y; y += 64) {   // four independent loads in flight
            const double v0 = part[(size_t)y * MP + col], v1 = part[(size_t)(y + 16) * MP + col];
            const double v2 = part[(size_t)(y + 32) * MP + col], v3 = part[(size_t)(y + 48) * MP + col];
            s += ((v0 + v1) + v2) + v3;
        }
        for (; y < Y; y += 16) s += part[(size_t)y * MP + col];
    }
    sh[g][c] = s;
    __syncthreads();
    if (g == 0 && col < MP) {
        double t = 0.0;
        for (int k = 0; k < 16; k++) t += sh[k][c];
        pos[col] = t;
    }
}

// Slots -> bins into acc (nbins complex).  An index i (a bin for the real
// kinds, M == nbins; a rotator position j for Shifting, M == L) owns the
// slots u = (i + sh) mod M + r*M, r < MP / M.  Shifting: bin b takes the
// positions j = i0 + k*N, i0 = (b - pidx) mod N, each rotated by
// rot[ridx + j].
__global__ __launch_bounds__(256) void k_pfinish(const double* __restrict__ pos, int MP, int M, int N, int shifting,
                                                 int sh, int pidx, int ridx, const float2* __restrict__ rot,
                                                 double* __restrict__ acc)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= N) return;
    const int R = MP / M;
    if (!shifting) {
        const int u = (b + sh) % M;
        double s = 0.0;
        for (int r = 0; r < R; r++) s += pos[(size_t)r * M + u];
        acc[2 * b] += s;
        return;
    }
    const int i0 = ((b - pidx) % N + N) % N;
    double re = 0.0, im = 0.0;
    for (int j = i0; j < M; j += N) {
        const int u = (j + sh) % M;
        double v = 0.0;
        for (int r = 0; r < R; r++) v += pos[(size_t)r * M + u];
        const float2 w = rot[ridx + j];
        re += v * (double)w.x;
        im += v * (double)w.y;
    }
    acc[2 * b] += re;
    acc[2 * b + 1] += im;
}

// The same map with one wave per bin, for bins that own many slots (short
// periods repeated R times, or Shifting's L/nbins positions per bin): lanes
// stride over the (position, repeat) pairs, then a fixed-order LDS tree.
__global__ __launch_bounds__(64) void k_pfinish_w(const double* __restrict__ pos, int MP, int M, int N, int shifting,
                                                  int sh, int pidx, int ridx, const float2* __restrict__ rot,
                                                  double* __restrict__ acc)
{
    __shared__ double s0[64], s1[64];
    const int b = blockIdx.x, lane = threadIdx.x;
    const int R = MP / M;
    const int i0 = shifting ? ((b - pidx) % N + N) % N : b;
    const int J = shifting ? (M - i0 + N - 1) / N : 1;   // positions j = i0 + k*N, k < J
    double re = 0.0, im = 0.0;
    for (long long t = lane; t < (long long)J * R; t += 64) {
        const int k = (int)(t / R), r = (int)(t % R);
        const int j = i0 + k * N;
        const double v = pos[(size_t)r * M + (j + sh) % M];
        if (shifting) {
            const float2 w = rot[ridx + j];
            re += v * (double)w.x;
            im += v * (double)w.y;
        } else {
            re += v;
        }
    }
    s0[lane] = re;
    s1[lane] = im;
    __syncthreads();
    for (int h = 32; h > 0; h >>= 1) {
        if (lane < h) {
            s0[lane] += s0[lane + h];
            s1[lane] += s1[lane + h];
        }
        __syncthreads();
    }
    if (lane == 0) {
        acc[2 * b] += s0[0];
        acc[2 * b + 1] += s1[0];
    }
}

__device__ inline void block_sum2(double& a, double& b, double (*sh)[256])
{
    sh[0][threadIdx.x] = a;
    sh[1][threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    a = sh[0][0];
    b = sh[1][0];
}

// tone t = (1/N) sum_b acc[b] e^{-2 pi i idx b / N}, idx = base + t * step
// (the forward DFT with 1/N of getFinalPCal at the bins it copies).
__global__ __launch_bounds__(256) void k_ptones(const double* __restrict__ acc, int N, long long base, long long step,
                                                double* __restrict__ tones)
{
    __shared__ double sh[2][256];
    const long long idx = base + (long long)blockIdx.x * step;
    double re = 0.0, im = 0.0;
    for (int b = threadIdx.x; b < N; b += 256) {
        const long long e = (idx * b) % N;
        double s, c;
        sincospi(2.0 * (double)e / (double)N, &s, &c);
        const double ar = acc[2 * b], ai = acc[2 * b + 1];
        re += ar * c + ai * s;
        im += ai * c - ar * s;
    }
    block_sum2(re, im, sh);
    if (threadIdx.x == 0) {
        tones[2 * blockIdx.x] = re / N;
        tones[2 * blockIdx.x + 1] = im / N;
    }
}

// |h(t_g)|, h(t) = sum_n T_n e^{+2 pi i n g / G}: the comb's time-domain
// response over one period on G points.
__global__ __launch_bounds__(256) void k_pdelay_mag(const double* __restrict__ tones, int nt, int G,
                                                    double* __restrict__ mag)
{
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= G) return;
    double hr = 0.0, hi = 0.0;
    for (int n = 0; n < nt; n++) {
        const long long e = ((long long)n * g) % G;
        double s, c;
        sincospi(2.0 * (double)e / (double)G, &s, &c);
        hr += tones[2 * n] * c - tones[2 * n + 1] * s;
        hi += tones[2 * n] * s + tones[2 * n + 1] * c;
    }
    mag[g] = sqrt(hr * hr + hi * hi);
}

// argmax of mag (lowest index on ties), parabolic refinement through the
// neighbours, delay in seconds in [0, 1/spacing).
__global__ __launch_bounds__(256) void k_pdelay_pick(const double* __restrict__ mag, int G, double spacing,
                                                     double* __restrict__ out)
{
    __shared__ double sv[256];
    __shared__ int si[256];
    double bv = -1.0;
    int bi = 0;
    for (int g = threadIdx.x; g < G; g += 256)
        if (mag[g] > bv) { bv = mag[g]; bi = g; }
    sv[threadIdx.x] = bv;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double v2 = sv[threadIdx.x + s];
            const int i2 = si[threadIdx.x + s];
            if (v2 > sv[threadIdx.x] || (v2 == sv[threadIdx.x] && i2 < si[threadIdx.x])) {
                sv[threadIdx.x] = v2;
                si[threadIdx.x] = i2;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int best = si[0];
        const double ym = mag[(best + G - 1) % G], y0 = mag[best], yp = mag[(best + 1) % G];
        const double den = ym - 2.0 * y0 + yp;
        const double d = (den != 0.0) ? 0.5 * (ym - yp) / den : 0.0;
        double t = ((double)best + d) / (double)G;
        t -= floor(t);
        out[0] = t / spacing;
    }
}

}  // namespace

struct swspec_hip_pcal {
    int kind = 0, device = 0;
    double fs = 0, spacing = 0;
    int offset_hz = 0;
    int nbins = 0, ntones = 0;
    long long L = 0;
    size_t pidx = 0, ridx = 0;
    uint64_t samplecount = 0;
    bool finalized = false;
    int nvalid = 0;               // tones getFinalPCal copies (before its break)
    long long tbase = 0, tstep = 1;
    double* d_acc = nullptr;      // nbins complex, FP64
    float2* d_rot = nullptr;      // 2L complex, FP32 (Shifting)
    double* d_part = nullptr;
    size_t part_cap = 0;
    double* d_pos = nullptr;
    size_t pos_cap = 0;
    double* d_tones = nullptr;    // ntones complex
    double* d_mag = nullptr;
    int mag_cap = 0;
    double* d_delay = nullptr;
    hipStream_t last = nullptr;   // stream of the last launch (buffers are reused across calls)
    std::vector<double> h_tones;
    std::string err;
};

namespace {

int pfail(swspec_hip_pcal* p, int code, const std::string& msg)
{
    if (p) p->err = msg;
    g_pcal_error = msg;
    return code;
}

#define PCHK(p, call)                                                                            \
    do {                                                                                         \
        hipError_t _e = (call);                                                                  \
        if (_e != hipSuccess)                                                                    \
            return pfail((p), _e == hipErrorOutOfMemory ? SWSPEC_ENOMEM : SWSPEC_EDEVICE,         \
                         std::string(#call) + ": " + hipGetErrorString(_e));                      \
    } while (0)

// PCal::gcd (PCal.cpp:115-127) on longs: double arguments truncate
long long pc_gcd(double da, double db)
{
    long long a = (long long)(long)da, b = (long long)(long)db;
    for (;;) {
        a = a % b;
        if (a == 0) return b;
        b = b % a;
        if (b == 0) return a;
    }
}

int grow(swspec_hip_pcal* p, double** buf, size_t* cap, size_t need)
{
    if (need <= *cap) return SWSPEC_OK;
    if (*buf) PCHK(p, hipFree(*buf));
    *buf = nullptr;
    *cap = 0;
    PCHK(p, hipMalloc(buf, need * sizeof(double)));
    *cap = need;
    return SWSPEC_OK;
}

}  // namespace

extern "C" int swspec_hip_pcal_create(double bandwidth_hz, double spacing_hz, int offset_hz, size_t sampleoffset,
                                      int kind, int device, swspec_hip_pcal** out)
{
    if (!out) return pfail(nullptr, SWSPEC_EINVAL, "null argument");
    *out = nullptr;
    if (!(bandwidth_hz > 0) || !(spacing_hz >= 1) || offset_hz < 0)
        return pfail(nullptr, SWSPEC_EINVAL, "bandwidth > 0, spacing >= 1 Hz and offset >= 0 required");
    if (kind < -1 || kind > SWSPEC_PCAL_IMPLICIT) return pfail(nullptr, SWSPEC_EINVAL, "unknown extractor kind");
    const double fs = 2 * bandwidth_hz;
    if (kind < 0) {   // PCal::getNew (:91-110)
        if (offset_hz == 0) {
            kind = SWSPEC_PCAL_TRIVIAL;
        } else {
            const int No = (int)(fs / (double)pc_gcd(offset_hz, fs));
            const int Np = (int)(fs / (double)pc_gcd(spacing_hz, fs));
            kind = (No % Np == 0) ? SWSPEC_PCAL_IMPLICIT : SWSPEC_PCAL_SHIFTING;
        }
    }
    if (kind != SWSPEC_PCAL_TRIVIAL && offset_hz == 0)
        return pfail(nullptr, SWSPEC_EINVAL, "Shifting/ImplicitShift need a non-zero offset");
    auto* p = new swspec_hip_pcal();
    p->kind = kind;
    p->device = device;
    p->fs = fs;
    double nb = 0, nt = 0, L = 0;
    if (kind == SWSPEC_PCAL_TRIVIAL) {   // :182-190, spacing arrives as int
        const int sp = (int)spacing_hz;
        p->spacing = sp;
        nb = fs / (double)pc_gcd(std::fabs((double)sp), fs);
        nt = std::floor(bandwidth_hz / sp);
    } else if (kind == SWSPEC_PCAL_SHIFTING) {   // :311-330
        p->spacing = spacing_hz;
        p->offset_hz = offset_hz;
        nb = fs / (double)pc_gcd(std::fabs(spacing_hz), fs);
        nt = std::floor((bandwidth_hz - offset_hz) / spacing_hz) + 1;
        L = fs / (double)pc_gcd(std::fabs((double)offset_hz), fs);
    } else {   // :522-528
        p->spacing = spacing_hz;
        p->offset_hz = offset_hz;
        nb = fs / (double)pc_gcd(std::fabs((double)offset_hz), fs);
        nt = std::floor((bandwidth_hz - offset_hz) / spacing_hz) + 1;
    }
    if (!(nb >= 1) || nb > (double)(1 << 28) || !(nt >= 1) || L > (double)(1 << 28)) {
        delete p;
        return pfail(nullptr, SWSPEC_EUNSUPPORTED, "tone count, bin count or rotator length out of range");
    }
    p->nbins = (int)nb;
    p->ntones = (int)nt;
    p->L = (long long)L;
    // tone bins getFinalPCal copies (:296-300, :503-512, :660-669)
    if (kind == SWSPEC_PCAL_TRIVIAL) {
        p->tbase = 0;
        p->tstep = 1;
        p->nvalid = std::min(p->ntones, p->nbins);
    } else {
        p->tstep = (long long)std::floor(p->nbins * (p->spacing / fs));
        if (kind == SWSPEC_PCAL_IMPLICIT) {
            // nbins * offset is an int product in the reference: wraps past 2^31
            const int32_t prod = (int32_t)((uint32_t)p->nbins * (uint32_t)offset_hz);
            p->tbase = (long long)std::floor((double)prod / fs);
        }
        p->nvalid = 0;
        for (int n = 0; n < p->ntones; n++) {
            const long long idx = p->tbase + (long long)n * p->tstep;
            if (idx < 0 || idx >= p->nbins) break;   // size_t index in the reference: negative is huge
            p->nvalid++;
        }
    }
    int rc;
    auto bail = [&](int r) { swspec_hip_pcal_destroy(p); return r; };
    if (hipSetDevice(device) != hipSuccess) return bail(pfail(nullptr, SWSPEC_EDEVICE, "hipSetDevice failed"));
    if (hipMalloc(&p->d_acc, sizeof(double) * 2 * p->nbins) != hipSuccess ||
        hipMalloc(&p->d_tones, sizeof(double) * 2 * std::max(p->ntones, 1)) != hipSuccess ||
        hipMalloc(&p->d_delay, sizeof(double)) != hipSuccess)
        return bail(pfail(nullptr, SWSPEC_ENOMEM, "device allocation failed"));
    if (kind == SWSPEC_PCAL_SHIFTING) {   // :353-359: FP32 values of FP64 cos/sin
        std::vector<float2> rot(2 * (size_t)p->L);
        const double dphi = 2 * M_PI * (-(double)offset_hz / fs);
        for (size_t n = 0; n < rot.size(); n++) {
            const double arg = dphi * (double)n;
            rot[n] = make_float2((float)std::cos(arg), (float)std::sin(arg));
        }
        if (hipMalloc(&p->d_rot, sizeof(float2) * rot.size()) != hipSuccess)
            return bail(pfail(nullptr, SWSPEC_ENOMEM, "device allocation failed"));
        if (hipMemcpy(p->d_rot, rot.data(), sizeof(float2) * rot.size(), hipMemcpyHostToDevice) != hipSuccess)
            return bail(pfail(nullptr, SWSPEC_EDEVICE, "rotator upload failed"));
    }
    rc = swspec_hip_pcal_clear(p);
    if (rc) return bail(rc);
    swspec_hip_pcal_adjust_offset(p, sampleoffset);
    *out = p;
    return SWSPEC_OK;
}

extern "C" void swspec_hip_pcal_destroy(swspec_hip_pcal* p)
{
    if (!p) return;
    (void)hipSetDevice(p->device);
    if (p->last) (void)hipStreamSynchronize(p->last);
    for (void* b : {(void*)p->d_acc, (void*)p->d_rot, (void*)p->d_part, (void*)p->d_pos, (void*)p->d_tones,
                    (void*)p->d_mag, (void*)p->d_delay})
        if (b) (void)hipFree(b);
    delete p;
}

extern "C" int swspec_hip_pcal_info(const swspec_hip_pcal* p, int* kind, int* ntones, int* nbins,
                                    int64_t* rotatorlen)
{
    if (!p) return pfail(nullptr, SWSPEC_EINVAL, "null extractor");
    if (kind) *kind = p->kind;
    if (ntones) *ntones = p->ntones;
    if (nbins) *nbins = p->nbins;
    if (rotatorlen) *rotatorlen = p->L;
    return SWSPEC_OK;
}

extern "C" int swspec_hip_pcal_clear(swspec_hip_pcal* p)
{
    if (!p) return pfail(nullptr, SWSPEC_EINVAL, "null extractor");
    PCHK(p, hipSetDevice(p->device));
    if (p->last) PCHK(p, hipStreamSynchronize(p->last));
    p->last = nullptr;
    p->samplecount = 0;
    p->finalized = false;
    PCHK(p, hipMemsetAsync(p->d_acc, 0, sizeof(double) * 2 * p->nbins, nullptr));
    PCHK(p, hipStreamSynchronize(nullptr));
    return SWSPEC_OK;
}

extern "C" int swspec_hip_pcal_adjust_offset(swspec_hip_pcal* p, size_t sampleoffset)
{
    if (!p) return pfail(nullptr, SWSPEC_EINVAL, "null extractor");
    p->pidx = sampleoffset % (size_t)p->nbins;
    p->ridx = p->kind == SWSPEC_PCAL_SHIFTING ? sampleoffset % (size_t)p->L : 0;
    return SWSPEC_OK;
}

extern "C" int swspec_hip_pcal_extract(swspec_hip_pcal* p, const float* d_samples, size_t len, void* stream)
{
    if (!p) return pfail(nullptr, SWSPEC_EINVAL, "null extractor");
    if (p->finalized) return pfail(p, SWSPEC_ESTATE, "extractAndIntegrate on a finalized extractor (clear first)");
    if (len == 0) return SWSPEC_OK;
    if (!d_samples) return pfail(p, SWSPEC_EINVAL, "null samples");
    if (len > (size_t)1 << 50) return pfail(p, SWSPEC_EINVAL, "length out of range");
    PCHK(p, hipSetDevice(p->device));
    hipStream_t st = (hipStream_t)stream;
    if (p->last && p->last != st) PCHK(p, hipStreamSynchronize(p->last));   // scratch reused across streams
    const bool shifting = p->kind == SWSPEC_PCAL_SHIFTING;
    const int N = p->nbins;
    const int M = shifting ? (int)p->L : N;
    const long long end = shifting ? (long long)(len - len % (size_t)M) : 0;
    const int jmax = shifting ? (M / N) * N : M;
    // slots per period MP = R * M: a multiple of 4, at least PF_MIN_MP, and
    // filling the last 256-slot tile (idle lanes cost bandwidth) as far as a
    // few more repeats allow
    const int R0 = std::max(1, (PF_MIN_MP + M - 1) / M);
    int R = 0;
    double best = 2.0;
    for (int r = R0; r <= 8 * R0 && (long long)M * r <= (1LL << 30); r++) {
        const long long mp = (long long)M * r;
        if (mp % 4) continue;
        const double waste = (double)(((mp / 4 + 63) / 64) * 256 - mp) / (double)mp;
        if (waste < best - 1e-12) { best = waste; R = r; }
        if (waste <= 1.0 / 64) break;
    }
    if (R == 0) return pfail(p, SWSPEC_EUNSUPPORTED, "no slot layout for this period");
    const int MP = M * R;
    const int a = (int)(((uintptr_t)d_samples / sizeof(float)) % 4);
    const int tiles = (MP / 4 + 63) / 64;
    const long long P = ((long long)len + a + MP - 1) / MP;
    long long Y = std::max<long long>(1, PF_TARGET_WG / tiles);   // one round of resident workgroups
    Y = std::min<long long>(Y, std::max<long long>(1, (P + PF_PMIN - 1) / PF_PMIN));
    Y = std::min<long long>(Y, std::max<long long>(1, (long long)(PF_PART_MAX / (size_t)MP)));
    Y = std::min<long long>(Y, 65535);
    const long long PG = (P + Y - 1) / Y;
    Y = (P + PG - 1) / PG;
    int rc = grow(p, &p->d_part, &p->part_cap, (size_t)Y * MP);
    if (rc) return rc;
    rc = grow(p, &p->d_pos, &p->pos_cap, (size_t)MP);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pfold, dim3(tiles, (unsigned)Y), dim3(PF_T), 0, st, d_samples, a, (long long)len, MP, M,
                       jmax, end, (int)PG, P, p->d_part);
    PCHK(p, hipGetLastError());
    hipLaunchKernelGGL(k_preduce, dim3((MP + 63) / 64), dim3(1024), 0, st, p->d_part, (int)Y, MP, p->d_pos);
    PCHK(p, hipGetLastError());
    // slot of bin b (real kinds): (b + a - pidx) mod M; of rotator position j (Shifting): (j + a) mod L
    const int sh = shifting ? a % M : (int)(((long long)a - (long long)p->pidx) % M + M) % M;
    const long long per_bin = (long long)R * (shifting ? (M + N - 1) / N : 1);
    if (per_bin >= 16)
        hipLaunchKernelGGL(k_pfinish_w, dim3(N), dim3(64), 0, st, p->d_pos, MP, M, N, shifting ? 1 : 0, sh,
                           (int)p->pidx, (int)p->ridx, p->d_rot, p->d_acc);
    else
        hipLaunchKernelGGL(k_pfinish, dim3((N + 255) / 256), dim3(256), 0, st, p->d_pos, MP, M, N,
                           shifting ? 1 : 0, sh, (int)p->pidx, (int)p->ridx, p->d_rot, p->d_acc);
    PCHK(p, hipGetLastError());
    p->last = st;
    // index updates of extractAndIntegrate (:286-289, :457-458, :632-635)
    if (shifting) {
        const size_t tail = len % (size_t)p->L;
        if (tail) {
            p->ridx = (p->ridx + tail) % (size_t)p->L;
            p->pidx = (p->pidx + tail % (size_t)N) % (size_t)N;
        }
    } else {
        p->pidx = (p->pidx + len % (size_t)N) % (size_t)N;
    }
    p->samplecount += len;
    return SWSPEC_OK;
}

extern "C" int swspec_hip_pcal_final(swspec_hip_pcal* p, float* out, double* out64, uint64_t* samplecount)
{
    if (!p) return pfail(nullptr, SWSPEC_EINVAL, "null extractor");
    PCHK(p, hipSetDevice(p->device));
    if (!p->finalized) {
        hipStream_t st = p->last;
        if (p->nvalid > 0) {
            hipLaunchKernelGGL(k_ptones, dim3(p->nvalid), dim3(256), 0, st, p->d_acc, p->nbins, p->tbase, p->tstep,
                               p->d_tones);
            PCHK(p, hipGetLastError());
        }
        PCHK(p, hipStreamSynchronize(st));
        p->h_tones.assign(2 * (size_t)std::max(p->nvalid, 1), 0.0);
        if (p->nvalid > 0)
            PCHK(p, hipMemcpy(p->h_tones.data(), p->d_tones, sizeof(double) * 2 * p->nvalid, hipMemcpyDeviceToHost));
        p->finalized = true;
        p->last = nullptr;
    }
    for (int n = 0; n < p->nvalid; n++) {
        if (out) {
            out[2 * n] = (float)p->h_tones[2 * n];
            out[2 * n + 1] = (float)p->h_tones[2 * n + 1];
        }
        if (out64) {
            out64[2 * n] = p->h_tones[2 * n];
            out64[2 * n + 1] = p->h_tones[2 * n + 1];
        }
    }
    if (samplecount) *samplecount = p->samplecount;
    return SWSPEC_OK;
}

extern "C" double swspec_hip_pcal_seconds(const swspec_hip_pcal* p)
{
    return (!p || p->fs == 0.0) ? 0.0 : (double)p->samplecount / p->fs;   // PCal.h getSeconds
}

extern "C" int swspec_hip_pcal_copied_tones(const swspec_hip_pcal* p)
{
    return p ? p->nvalid : 0;
}

extern "C" int swspec_hip_pcal_delay(swspec_hip_pcal* p, int upsample, double* seconds)
{
    if (!p || !seconds) return pfail(p, SWSPEC_EINVAL, "null argument");
    if (!p->finalized) return pfail(p, SWSPEC_ESTATE, "delay needs getFinalPCal first");
    if (upsample < 2 || upsample > 4096) return pfail(p, SWSPEC_EINVAL, "upsample must be 2..4096");
    if (p->nvalid < 2) return pfail(p, SWSPEC_EINVAL, "delay needs at least two tones");
    const long long G = (long long)upsample * p->nvalid;
    if (G > (1 << 24)) return pfail(p, SWSPEC_EINVAL, "upsample * tones too large");
    PCHK(p, hipSetDevice(p->device));
    if (G > p->mag_cap) {
        if (p->d_mag) PCHK(p, hipFree(p->d_mag));
        p->d_mag = nullptr;
        p->mag_cap = 0;
        PCHK(p, hipMalloc(&p->d_mag, sizeof(double) * G));
        p->mag_cap = (int)G;
    }
    hipLaunchKernelGGL(k_pdelay_mag, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, nullptr, p->d_tones,
                       p->nvalid, (int)G, p->d_mag);
    PCHK(p, hipGetLastError());
    hipLaunchKernelGGL(k_pdelay_pick, dim3(1), dim3(256), 0, nullptr, p->d_mag, (int)G, p->spacing, p->d_delay);
    PCHK(p, hipGetLastError());
    PCHK(p, hipMemcpy(seconds, p->d_delay, sizeof(double), hipMemcpyDeviceToHost));
    return SWSPEC_OK;
}

extern "C" const char* swspec_hip_pcal_last_error(const swspec_hip_pcal* p)
{
    return p ? p->err.c_str() : g_pcal_error.c_str();
}
