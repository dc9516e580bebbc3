/*
 * swspec_oracle.c -- CPU restatement of the swspec hot path (checker only).
 *
 * TEST INFRASTRUCTURE: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py only.  Nothing in the product links it.
 *
 * Parity pins (what checks each part):
 *   so_derive          hand-computed configuration tables (tests/test_oracle_settings.py)
 *   so_unpack          exhaustive closed-form byte tables + reference LUT literals
 *   so_window_table    closed forms of the reference source; IPP windows only
 *                      against Intel's published formula (no IPP available)
 *   so_rfft64/32       brute-force DFT, numpy pocketfft, Parseval
 *   so_domaths         tone-on-bin and 1/37-period known answers, combine
 *                      identity, FP32-vs-FP64 agreement, XY* of equal inputs
 *   so_extract_pcal*   brute-force per-sample rotator/fold (PCal.cpp:139-175)
 *   so_pcal_tones64    PCal.cpp unit-test KAT (phase -90 + 5*t deg within 0.1 deg)
 *   so_pcalx_*         the same unit test per extractor class, the brute-force
 *                      rotating fold, the DFT closed form, impulse responses
 *   so_pcal_delay64    delayed synthetic combs (no reference code exists)
 *   so_mk5b_strip      synthetic frames with known payload
 */
#include "swspec_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------- */
/* settings derivation                                                       */
/* ------------------------------------------------------------------------- */

/* src/Helpers.cpp:150-162 */
long long so_gcd(long long a, long long b)
{
    for (;;) {
        a = a % b;
        if (a == 0) return b;
        b = b % a;
        if (b == 0) return a;
    }
}

/* src/swspectrometer.cpp:67-231 */
int so_derive(const so_ini* ini, so_settings* s)
{
    memset(s, 0, sizeof(*s));
    s->num_cores = ini->num_cores > 0 ? ini->num_cores : 1;
    s->vlba_fanout = ini->vlba_fanout;
    /* :70 default, :104-106 INI key (MByte, double-buffered: *1024*1024/2) */
    if (ini->max_source_buffer_mb >= 0) {
        s->max_rawbuf_size = (size_t)ini->max_source_buffer_mb;
        s->max_rawbuf_size *= 1024 * 1024 / 2;
    } else {
        s->max_rawbuf_size = (size_t)ini->platform_max_raw_mb * 1024 * 1024;
    }
    s->fft_points = ini->fft_points;
    s->fft_integ_seconds = ini->fft_integ_seconds;
    s->fft_overlap_factor = ini->fft_overlap_factor;
    s->window = ini->window;
    s->samplingfreq = (float)((double)ini->bandwidth_hz * 2.0);  /* :116 (float *= double) */
    s->pcaloffsethz = ini->pcal_offset_hz;
    s->pcalharmonicshz = ini->pcal_harmonics_hz;
    s->source_format = ini->source_format;
    s->bits_per_sample = ini->bits_per_sample;
    s->channelorder_increasing = ini->channelorder_increasing;
    s->source_channels = ini->source_channels;
    s->extract_pcal = ini->extract_pcal;
    s->num_sources = ini->num_sources;
    s->num_xpols = (ini->do_xpol && ini->num_sources == 2) ? 1 : 0; /* :156-158, :342-356 */

    /* :148-155 channel checks, then :162-163 decrement */
    if (ini->use_channel_file1 < 1 || ini->use_channel_file1 > ini->source_channels) return -1;
    if (ini->use_channel_file2 < 1 || ini->use_channel_file2 > ini->source_channels) return -1;
    s->use_channel_file1 = ini->use_channel_file1 - 1;
    s->use_channel_file2 = ini->use_channel_file2 - 1;

    /* :165-176.  Types follow Settings.h: samplingfreq and fft_integ_seconds are
     * float, fft_points is size_t, so (fs*T)/N is evaluated in float. */
    s->dt = (float)(1.0 / s->samplingfreq);
    s->df = s->samplingfreq / (float)s->fft_points;
    {
        float prod = s->samplingfreq * s->fft_integ_seconds;
        float q = prod / (float)(size_t)s->fft_points;
        s->averaged_ffts = (int)q;
    }
    s->fft_overlap_points = (int)((size_t)s->fft_points / (size_t)s->fft_overlap_factor);
    s->fft_ssb_points = (int)((size_t)s->fft_points / 2 + 1);
    s->rawbytes_per_channelsample = (s->bits_per_sample * s->source_channels) / 8.0;
    s->raw_fullfft_bytes = (size_t)(int)((double)s->fft_points * s->rawbytes_per_channelsample);
    s->raw_overlap_bytes = (size_t)(int)((double)s->fft_overlap_points * s->rawbytes_per_channelsample);
    s->fft_bytes = (size_t)s->fft_points * sizeof(float);
    s->fft_bytes_ssb = (size_t)s->fft_ssb_points * sizeof(float);
    s->fft_bytes_xpol = 2 * s->fft_bytes_ssb;

    /* :178-196 */
    if (s->extract_pcal) {
        if ((long long)s->pcaloffsethz == 0) return -2;    /* a % 0 in Helpers::gcd */
        s->pcal_tonebins = (int)(s->samplingfreq / s->pcalharmonicshz);
        {
            long long g = so_gcd((long long)s->samplingfreq, (long long)s->pcaloffsethz);
            s->pcal_rotatorlen = (int)(s->samplingfreq / (float)g);
        }
        s->pcal_pulses = s->pcal_rotatorlen / s->pcal_tonebins;
        s->pcal_result_bytes = (size_t)s->pcal_tonebins * 2 * sizeof(float);
        s->pcal_pulses_per_fft = (int)((size_t)s->fft_points / (size_t)s->pcal_rotatorlen);
        if (((size_t)s->fft_points % (size_t)s->pcal_rotatorlen) != 0) {
            s->extract_pcal = 0;
        }
    } else {
        s->pcal_tonebins = 0;
        s->pcal_rotatorlen = 0;
        s->pcal_pulses = 0;
        s->pcal_result_bytes = 0;
        s->pcal_pulses_per_fft = 0;
    }

    /* :198-227 */
    s->max_rawbuf_size = s->max_rawbuf_size / (size_t)s->num_cores;
    {
        size_t tent = s->raw_fullfft_bytes * (size_t)s->averaged_ffts;
        if (tent > s->max_rawbuf_size) {
            s->max_spectra_per_buffer = 0;
            s->max_specffts_per_buffer = 1;
            s->max_buffers_per_spectrum = s->averaged_ffts;
            tent = s->raw_fullfft_bytes;
            while ((tent <= s->max_rawbuf_size) && ((s->max_buffers_per_spectrum % 2) == 0)) {
                tent *= 2;
                s->max_buffers_per_spectrum /= 2;
                s->max_specffts_per_buffer *= 2;
            }
        } else {
            s->max_spectra_per_buffer = 1;
            s->max_specffts_per_buffer = s->averaged_ffts;
            s->max_buffers_per_spectrum = 0;
            while ((2 * tent) <= s->max_rawbuf_size) {
                tent *= 2;
                s->max_spectra_per_buffer *= 2;
            }
        }
        s->rawbuf_size = tent;
    }
    /* :230-231 */
    s->core_averaged_ffts = s->max_specffts_per_buffer;
    s->core_overlapped_ffts = s->fft_overlap_factor * s->max_specffts_per_buffer - (s->fft_overlap_factor - 1);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* decoders                                                                  */
/* ------------------------------------------------------------------------- */

/* map_rev of DataUnpackers.cpp:212 (float literal 3.3359 rounded to float) */
const float so_map_rev[4] = { 1.0f, -1.0f, 3.3359f, -3.3359f };

/* DataUnpackerFactory.h:38-108 restricted to the 2-bit decoders */
int so_pick_unpacker(const so_settings* s)
{
    if (s->source_format == SO_FMT_RAWSIGNED || s->source_format == SO_FMT_RAWUNSIGNED) {
        if (s->bits_per_sample == 2 && (s->source_channels % 4) == 0) return 1; /* :246-249 */
        if (s->bits_per_sample == 2 && s->source_channels == 1) return 2;       /* :307-310 */
        return 0;
    }
    if (s->source_format == SO_FMT_MK5B) {
        /* Mk5BUnpacker::canHandleConfig (:444-447) accepts any even channel
         * count; only its 4- and 8-channel branches are well defined
         * (the 2-channel branch indexes a 256-entry LUT with 16-bit words and
         * the 16-channel branch prints every sample). */
        if (s->bits_per_sample == 2 && (s->source_channels == 4 || s->source_channels == 8)) return 3;
        return 0;
    }
    if (s->source_format == SO_FMT_VLBA) {
        /* VLBAUnpacker::canHandleConfig (:530-533); restated for the modes
         * with one byte per time sample: 4 channels x 2 bits or 8 x 1 bit,
         * fanout 1, 2 or 4 */
        const int f = s->vlba_fanout;
        if (f != 1 && f != 2 && f != 4) return 0;
        if ((s->bits_per_sample == 2 && s->source_channels == 4) ||
            (s->bits_per_sample == 1 && s->source_channels == 8)) return 4;
        return 0;
    }
    if (s->source_format == SO_FMT_MARK5B) {
        /* Mark5BUnpacker::canHandleConfig (:755-758): 1 or 2 bits, even
         * channel count; mark5access takes 1..32 bitstreams (decimation 1) */
        const int B = s->source_channels * s->bits_per_sample;
        if ((s->bits_per_sample == 1 || s->bits_per_sample == 2) && (s->source_channels % 2) == 0 &&
            (B == 2 || B == 4 || B == 8 || B == 16 || B == 32)) return 5;
        return 0;
    }
    return 0;
}

size_t so_unpack(const so_settings* s, const uint8_t* src, float* dst, size_t count, int channel)
{
    int kind = so_pick_unpacker(s);
    size_t smp;
    if (kind == 1 || kind == 3) {
        /* TwoBitUnpacker (:191-244); Mk5BUnpacker 4/8-ch (:358-407) is the
         * same with increasing channel order always. */
        int incr = (kind == 3) ? 1 : s->channelorder_increasing;
        int step = s->source_channels / 4;
        unsigned shift = incr ? (unsigned)(2 * (channel % 4)) : (unsigned)(6 - 2 * (channel % 4));
        const uint8_t* p = incr ? src + channel / 4 : src + (step - 1) - channel / 4;
        for (smp = 0; smp < count; smp += 8) {
            for (int off = 0; off < 8; off++) {
                dst[smp + off] = so_map_rev[(p[off * step] >> shift) & 3u];
            }
            p += 8 * step;
        }
        return smp;
    }
    if (kind == 2) {
        /* TwoBitSinglechannelUnpacker (:261-305), documented layout (:267-268):
         * four samples per byte, oldest sample in the MSB. */
        for (smp = 0; smp < count; smp += 16) {
            for (int off = 0; off < 4; off++) {
                uint8_t v = src[off];
                for (int k = 0; k < 4; k++) dst[smp + 4 * off + k] = so_map_rev[(v >> (2 * (3 - k))) & 3u];
            }
            src += 4;
        }
        return smp;
    }
    if (kind == 4) return so_vlba_unpack_mode(src, (long)count, count, channel, s->vlba_fanout, s->bits_per_sample, dst);
    if (kind == 5) {
        const long bytes = (long)((count * (size_t)s->source_channels * (size_t)s->bits_per_sample + 7) / 8);
        return so_mark5b_unpack(src, bytes, count, channel, s->source_channels, s->bits_per_sample, dst);
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* windows (TaskCoreIPP.cpp:569-605)                                          */
/* ------------------------------------------------------------------------- */

static double win_value(int type, long i, long points)
{
    double n1 = (double)(points - 1);
    switch (type) {
    case SO_WIN_NONE:
        return 1.0;
    case SO_WIN_COSINE:   /* :577-579: Ipp32f(i) - 0.5*Ipp32f(points) + 0.5, in double */
        return cos(M_PI * ((double)(float)i - 0.5 * (double)(float)points + 0.5) / (double)(float)points);
    case SO_WIN_COSINE2: {/* :581-586 */
        double w = cos(M_PI * ((double)(float)i - 0.5 * (double)(float)points + 0.5) / (double)(float)points);
        return w * w;
    }
    case SO_WIN_HAMMING:  /* ippsWinHamming: 0.54 - 0.46 cos(2 pi n/(N-1)) */
        return 0.54 - 0.46 * cos(2.0 * M_PI * (double)i / n1);
    case SO_WIN_HANN:     /* ippsWinHann: 0.5 - 0.5 cos(2 pi n/(N-1)) */
        return 0.5 - 0.5 * cos(2.0 * M_PI * (double)i / n1);
    case SO_WIN_BLACKMAN: {/* ippsWinBlackman(alpha): (a+1)/2 - 0.5cos(x) - (a/2)cos(2x), a = 0.5 (:593) */
        double a = 0.5, x = 2.0 * M_PI * (double)i / n1;
        return (a + 1.0) / 2.0 - 0.5 * cos(x) - (a / 2.0) * cos(2.0 * x);
    }
    default:
        return 1.0;
    }
}

void so_window_table(int type, long points, float* out)
{
    for (long i = 0; i < points; i++) out[i] = (float)win_value(type, i, points);
}

void so_window_table64(int type, long points, double* out)
{
    for (long i = 0; i < points; i++) out[i] = win_value(type, i, points);
}

/* ------------------------------------------------------------------------- */
/* mixed-radix Stockham FFT, FP64 and FP32 (twiddles always from FP64)        */
/* ------------------------------------------------------------------------- */

#define SO_MAXF 64

typedef struct fplan {
    long n;
    int nf;
    int fac[SO_MAXF];
    double* tw64;   /* concatenated per-stage twiddles */
    float* tw32;
    long twoff[SO_MAXF];
    double* rtw64;  /* r2c split twiddles W_{2n}^k, k < n */
    float* rtw32;
} fplan;

static int factorize(long n, int* fac)
{
    int nf = 0;
    while (n % 4 == 0 && nf < SO_MAXF) { fac[nf++] = 4; n /= 4; }
    while (n % 2 == 0 && nf < SO_MAXF) { fac[nf++] = 2; n /= 2; }
    for (long p = 3; n > 1 && nf < SO_MAXF; p += 2) {
        while (n % p == 0 && nf < SO_MAXF) { fac[nf++] = (int)p; n /= p; }
        if (p > 4096) return -1;
    }
    return n == 1 ? nf : -1;
}

static fplan* plan_new(long n, int real_split)
{
    fplan* p = (fplan*)calloc(1, sizeof(fplan));
    p->n = n;
    p->nf = factorize(n, p->fac);
    if (p->nf < 0) { free(p); return NULL; }
    long total = 0, ns = 1;
    for (int s = 0; s < p->nf; s++) { p->twoff[s] = total; total += ns * (p->fac[s] - 1); ns *= p->fac[s]; }
    p->tw64 = (double*)malloc(sizeof(double) * 2 * (size_t)(total + 1));
    p->tw32 = (float*)malloc(sizeof(float) * 2 * (size_t)(total + 1));
    ns = 1;
    for (int s = 0; s < p->nf; s++) {
        int R = p->fac[s];
        for (long k = 0; k < ns; k++)
            for (int r = 1; r < R; r++) {
                /* W_{ns R}^{k r}, angle reduced exactly in integers */
                long long e = ((long long)k * r) % (ns * R);
                double a = -2.0 * M_PI * (double)e / (double)(ns * R);
                long idx = p->twoff[s] + k * (R - 1) + (r - 1);
                p->tw64[2 * idx] = cos(a);
                p->tw64[2 * idx + 1] = sin(a);
                p->tw32[2 * idx] = (float)cos(a);
                p->tw32[2 * idx + 1] = (float)sin(a);
            }
        ns *= R;
    }
    if (real_split) {
        p->rtw64 = (double*)malloc(sizeof(double) * 2 * (size_t)n);
        p->rtw32 = (float*)malloc(sizeof(float) * 2 * (size_t)n);
        for (long k = 0; k < n; k++) {
            double a = -2.0 * M_PI * (double)k / (double)(2 * n);
            p->rtw64[2 * k] = cos(a);
            p->rtw64[2 * k + 1] = sin(a);
            p->rtw32[2 * k] = (float)cos(a);
            p->rtw32[2 * k + 1] = (float)sin(a);
        }
    }
    return p;
}

static void plan_free(fplan* p)
{
    if (!p) return;
    free(p->tw64); free(p->tw32); free(p->rtw64); free(p->rtw32); free(p);
}

/* One generic body for both precisions. */
#define SO_FFT_BODY(T, TW)                                                           \
    long n = p->n;                                                                   \
    T* in = data;                                                                    \
    T* out = work;                                                                   \
    long ns = 1;                                                                     \
    for (int s = 0; s < p->nf; s++) {                                                \
        int R = p->fac[s];                                                           \
        long nr = n / R;                                                             \
        const T* tw = p->TW + 2 * p->twoff[s];                                       \
        T* rootr = (T*)malloc(sizeof(T) * 2 * (size_t)R); T* rooti = rootr + R;      \
        for (int e = 0; e < R; e++) {                                                \
            rootr[e] = (T)cos(-2.0 * M_PI * e / R);                                  \
            rooti[e] = (T)sin(-2.0 * M_PI * e / R);                                  \
        }                                                                            \
        for (long j = 0; j < nr; j++) {                                              \
            long k = j % ns;                                                         \
            T vr[64], vi[64];                                                        \
            T* vre = vr; T* vim = vi;                                                \
            T* bigr = NULL; T* bigi = NULL;                                          \
            if (R > 64) { bigr = (T*)malloc(sizeof(T) * R * 2); bigi = bigr + R;     \
                          vre = bigr; vim = bigi; }                                  \
            for (int r = 0; r < R; r++) {                                            \
                T xr = in[2 * (j + r * nr)], xi = in[2 * (j + r * nr) + 1];          \
                if (r > 0) {                                                         \
                    T wr = tw[2 * (k * (R - 1) + r - 1)];                            \
                    T wi = tw[2 * (k * (R - 1) + r - 1) + 1];                        \
                    T tr = xr * wr - xi * wi; xi = xr * wi + xi * wr; xr = tr;       \
                }                                                                    \
                vre[r] = xr; vim[r] = xi;                                            \
            }                                                                        \
            long od = (j / ns) * ns * R + k;                                         \
            if (R == 2) {                                                            \
                T ar = vre[0] + vre[1], ai = vim[0] + vim[1];                        \
                T br = vre[0] - vre[1], bi = vim[0] - vim[1];                        \
                out[2 * od] = ar; out[2 * od + 1] = ai;                              \
                out[2 * (od + ns)] = br; out[2 * (od + ns) + 1] = bi;                \
            } else if (R == 4) {                                                     \
                T s0r = vre[0] + vre[2], s0i = vim[0] + vim[2];                      \
                T d0r = vre[0] - vre[2], d0i = vim[0] - vim[2];                      \
                T s1r = vre[1] + vre[3], s1i = vim[1] + vim[3];                      \
                T d1r = vre[1] - vre[3], d1i = vim[1] - vim[3];                      \
                out[2 * od] = s0r + s1r; out[2 * od + 1] = s0i + s1i;                \
                out[2 * (od + 2 * ns)] = s0r - s1r;                                  \
                out[2 * (od + 2 * ns) + 1] = s0i - s1i;                              \
                /* -i * d1 */                                                        \
                out[2 * (od + ns)] = d0r + d1i; out[2 * (od + ns) + 1] = d0i - d1r;  \
                out[2 * (od + 3 * ns)] = d0r - d1i;                                  \
                out[2 * (od + 3 * ns) + 1] = d0i + d1r;                              \
            } else if (R == 3) {                                                     \
                const T s3 = (T)0.86602540378443864676;                              \
                T tr = vre[1] + vre[2], ti = vim[1] + vim[2];                        \
                T dr = (vre[1] - vre[2]) * s3, di = (vim[1] - vim[2]) * s3;          \
                T ar = vre[0] - (T)0.5 * tr, ai = vim[0] - (T)0.5 * ti;              \
                out[2 * od] = vre[0] + tr; out[2 * od + 1] = vim[0] + ti;            \
                out[2 * (od + ns)] = ar + di; out[2 * (od + ns) + 1] = ai - dr;      \
                out[2 * (od + 2 * ns)] = ar - di;                                    \
                out[2 * (od + 2 * ns) + 1] = ai + dr;                                \
            } else if (R == 5) {                                                     \
                const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410; \
                const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;  \
                T t1r = vre[1] + vre[4], t1i = vim[1] + vim[4];                      \
                T t2r = vre[2] + vre[3], t2i = vim[2] + vim[3];                      \
                T t3r = vre[1] - vre[4], t3i = vim[1] - vim[4];                      \
                T t4r = vre[2] - vre[3], t4i = vim[2] - vim[3];                      \
                T a1r = vre[0] + c1 * t1r + c2 * t2r, a1i = vim[0] + c1 * t1i + c2 * t2i; \
                T a2r = vre[0] + c2 * t1r + c1 * t2r, a2i = vim[0] + c2 * t1i + c1 * t2i; \
                T b1r = s1 * t3r + s2 * t4r, b1i = s1 * t3i + s2 * t4i;              \
                T b2r = s2 * t3r - s1 * t4r, b2i = s2 * t3i - s1 * t4i;              \
                out[2 * od] = vre[0] + t1r + t2r;                                    \
                out[2 * od + 1] = vim[0] + t1i + t2i;                                \
                out[2 * (od + ns)] = a1r + b1i; out[2 * (od + ns) + 1] = a1i - b1r;  \
                out[2 * (od + 4 * ns)] = a1r - b1i;                                  \
                out[2 * (od + 4 * ns) + 1] = a1i + b1r;                              \
                out[2 * (od + 2 * ns)] = a2r + b2i;                                  \
                out[2 * (od + 2 * ns) + 1] = a2i - b2r;                              \
                out[2 * (od + 3 * ns)] = a2r - b2i;                                  \
                out[2 * (od + 3 * ns) + 1] = a2i + b2r;                              \
            } else {                                                                 \
                for (int m = 0; m < R; m++) {                                        \
                    T ar = 0, ai = 0;                                                \
                    for (int r = 0; r < R; r++) {                                    \
                        long e = ((long)m * r) % R;                                  \
                        ar += vre[r] * rootr[e] - vim[r] * rooti[e];                 \
                        ai += vre[r] * rooti[e] + vim[r] * rootr[e];                 \
                    }                                                                \
                    out[2 * (od + m * ns)] = ar; out[2 * (od + m * ns) + 1] = ai;    \
                }                                                                    \
            }                                                                        \
            if (bigr) free(bigr);                                                    \
        }                                                                            \
        free(rootr);                                                                 \
        T* t = in; in = out; out = t;                                                \
        ns *= R;                                                                     \
    }                                                                                \
    if (in != data) memcpy(data, in, sizeof(T) * 2 * (size_t)n);


static void cfft_run64(const fplan* p, double* data, double* work) { SO_FFT_BODY(double, tw64) }
static void cfft_run32(const fplan* p, float* data, float* work) { SO_FFT_BODY(float, tw32) }

int so_cfft64(double* z, long n)
{
    fplan* p = plan_new(n, 0);
    if (!p) return -1;
    double* w = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    cfft_run64(p, z, w);
    free(w);
    plan_free(p);
    return 0;
}

/* r2c split: z = x[2m] + i x[2m+1], Z = FFT_M(z), X[k] = E[k] + W_N^k O[k]. */
#define SO_RSPLIT(T, RTW)                                                              \
    for (long k = 0; k <= m; k++) {                                                    \
        long km = (k == 0 || k == m) ? 0 : m - k;                                      \
        long kk = (k == m) ? 0 : k;                                                    \
        T zr = z[2 * kk], zi = z[2 * kk + 1];                                          \
        T cr = z[2 * km], ci = -z[2 * km + 1];                                         \
        T er = (T)0.5 * (zr + cr), ei = (T)0.5 * (zi + ci);                            \
        /* O = (Z - conj(Zm)) / (2i) = -i/2 (Z - conj(Zm)) */                          \
        T dr = zr - cr, di = zi - ci;                                                  \
        T orr = (T)0.5 * di, oi = (T)-0.5 * dr;                                        \
        T wr, wi;                                                                      \
        if (k == m) { wr = -1; wi = 0; } else { wr = p->RTW[2 * k]; wi = p->RTW[2 * k + 1]; } \
        X[2 * k] = er + (orr * wr - oi * wi);                                          \
        X[2 * k + 1] = ei + (orr * wi + oi * wr);                                      \
    }                                                                                  \
    X[1] = 0; X[2 * m + 1] = 0;

int so_rfft64(const double* x, long n, double* X)
{
    if (n % 2) return -1;
    long m = n / 2;
    fplan* p = plan_new(m, 1);
    if (!p) return -1;
    double* z = (double*)malloc(sizeof(double) * 4 * (size_t)m);
    double* w = z + 2 * m;
    memcpy(z, x, sizeof(double) * (size_t)n);
    cfft_run64(p, z, w);
    SO_RSPLIT(double, rtw64)
    free(z);
    plan_free(p);
    return 0;
}

static void rfft32_plan(const fplan* p, const float* x, float* X, float* z /* 4m floats */)
{
    long m = p->n;
    float* w = z + 2 * m;
    memcpy(z, x, sizeof(float) * 2 * (size_t)m);
    cfft_run32(p, z, w);
    SO_RSPLIT(float, rtw32)
}

static void rfft64_plan(const fplan* p, const double* x, double* X, double* z)
{
    long m = p->n;
    double* w = z + 2 * m;
    memcpy(z, x, sizeof(double) * 2 * (size_t)m);
    cfft_run64(p, z, w);
    SO_RSPLIT(double, rtw64)
}

int so_rfft32(const float* x, long n, float* X)
{
    if (n % 2) return -1;
    fplan* p = plan_new(n / 2, 1);
    if (!p) return -1;
    float* z = (float*)malloc(sizeof(float) * 2 * (size_t)n);
    rfft32_plan(p, x, X, z);
    free(z);
    plan_free(p);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* phase-cal                                                                 */
/* ------------------------------------------------------------------------- */

/* TaskCoreIPP.cpp:129-136 (rotator, FP64 -> FP32) and :633-693 (fold) */
void so_extract_pcal32(const so_settings* s, const float* x, long points, float* pcal)
{
    int L = s->pcal_rotatorlen, T = s->pcal_tonebins;
    float* rot = (float*)malloc(sizeof(float) * 2 * (size_t)L);
    double dphi = 2 * M_PI * (-(double)s->pcaloffsethz / (double)s->samplingfreq);
    for (int i = 0; i < L; i++) {
        double arg = dphi * (double)i;
        rot[2 * i] = (float)cos(arg);
        rot[2 * i + 1] = (float)sin(arg);
    }
    for (long n = 0; n < points; n += L) {
        /* ippsMul_32f32fc then ippsAdd_32fc_I per pulse */
        for (int p = 0; p < L / T; p++)
            for (int b = 0; b < T; b++) {
                long i = (long)p * T + b;
                float xr = x[n + i] * rot[2 * i];
                float xi = x[n + i] * rot[2 * i + 1];
                pcal[2 * b] += xr;
                pcal[2 * b + 1] += xi;
            }
    }
    free(rot);
}

void so_extract_pcal64(const so_settings* s, const double* x, long points, double* pcal)
{
    so_pcal_fold64(x, points, (double)s->samplingfreq, (double)s->pcaloffsethz, s->pcal_tonebins,
                   s->pcal_rotatorlen, 0, 0, pcal);
}

void so_pcal_fold64(const double* x, long len, double fs_hz, double offset_hz, int nbins,
                    long rotatorlen, uint64_t sampleoffset, int rot32, double* acc)
{
    double dphi = 2 * M_PI * (-offset_hz / fs_hz);
    double* rc = (double*)malloc(sizeof(double) * 2 * (size_t)rotatorlen);
    for (long i = 0; i < rotatorlen; i++) {
        double arg = dphi * (double)i;
        rc[2 * i] = rot32 ? (double)(float)cos(arg) : cos(arg);
        rc[2 * i + 1] = rot32 ? (double)(float)sin(arg) : sin(arg);
    }
    for (long n = 0; n < len; n++) {
        uint64_t no = sampleoffset + (uint64_t)n;
        long i = (long)(no % (uint64_t)rotatorlen);
        long b = (long)(no % (uint64_t)nbins);
        acc[2 * b] += x[n] * rc[2 * i];
        acc[2 * b + 1] += x[n] * rc[2 * i + 1];
    }
    free(rc);
}

/* PCal.cpp:328 */
int so_pcal_ntones(double bandwidth_hz, double spacing_hz, double offset_hz)
{
    return (int)floor((bandwidth_hz - offset_hz) / spacing_hz) + 1;
}

/* PCal.cpp:489-515: forward DFT scaled 1/nbins, bins n*floor(nbins*spacing/fs). */
void so_pcal_tones64(const double* folded, int nbins, int ntones, double spacing_hz,
                     double fs_hz, double* tones_out)
{
    size_t step = (size_t)floor(nbins * (spacing_hz / fs_hz));
    for (int t = 0; t < ntones; t++) {
        size_t idx = (size_t)t * step;
        tones_out[2 * t] = 0; tones_out[2 * t + 1] = 0;
        if (idx >= (size_t)nbins) continue;
        double ar = 0, ai = 0;
        for (int b = 0; b < nbins; b++) {
            long long e = ((long long)idx * b) % nbins;
            double a = -2.0 * M_PI * (double)e / (double)nbins;
            ar += folded[2 * b] * cos(a) - folded[2 * b + 1] * sin(a);
            ai += folded[2 * b] * sin(a) + folded[2 * b + 1] * cos(a);
        }
        tones_out[2 * t] = ar / nbins;
        tones_out[2 * t + 1] = ai / nbins;
    }
}

/* PCal.cpp:139-175 in FP64 throughout (the reference accumulates in FP32). */
void so_pcal_reference64(const double* x, long len, double bandwidth_hz, double spacing_hz,
                         double offset_hz, uint64_t sampleoffset, double* tones_out)
{
    double fs = 2 * bandwidth_hz;
    int ntones = so_pcal_ntones(bandwidth_hz, spacing_hz, offset_hz);
    long nbins = 2 * ntones;
    long long period = (long long)fs / so_gcd((long long)offset_hz, (long long)fs);
    double dphi = 2 * M_PI * (-offset_hz / fs);
    double* acc = (double*)calloc((size_t)nbins * 2, sizeof(double));
    for (long n = 0; n < len; n++) {
        uint64_t no = (uint64_t)n + sampleoffset;
        long bin = (long)(no % (uint64_t)nbins);
        double phi = dphi * (double)(no % (uint64_t)period);
        acc[2 * bin] += cos(phi) * x[n];
        acc[2 * bin + 1] += sin(phi) * x[n];
    }
    /* the reference DFT is forward with 1/N and copies the first ntones bins */
    for (int t = 0; t < ntones; t++) {
        double ar = 0, ai = 0;
        for (long b = 0; b < nbins; b++) {
            long long e = ((long long)t * b) % nbins;
            double a = -2.0 * M_PI * (double)e / (double)nbins;
            ar += acc[2 * b] * cos(a) - acc[2 * b + 1] * sin(a);
            ai += acc[2 * b] * sin(a) + acc[2 * b + 1] * cos(a);
        }
        tones_out[2 * t] = ar / nbins;
        tones_out[2 * t + 1] = ai / nbins;
    }
    free(acc);
}

/* ------------------------------------------------------------------------- */
/* PCal extractor classes (src/IA-32/PhaseCal/PCal.cpp)                      */
/* ------------------------------------------------------------------------- */

/* PCal::gcd (:115-127) takes longs: double arguments truncate */
static long long pc_gcd(double a, double b) { return so_gcd((long long)(long)a, (long long)(long)b); }

int so_pcalx_init(so_pcalx* p, int kind, double bw, double spacing, int offset, size_t sampleoffset)
{
    memset(p, 0, sizeof(*p));
    double fs = 2 * bw;
    if (kind < 0) {   /* PCal::getNew (:91-110) */
        if (offset == 0) {
            kind = SO_PCAL_TRIVIAL;
        } else {
            int No = (int)(fs / (double)pc_gcd(offset, fs));
            int Np = (int)(fs / (double)pc_gcd(spacing, fs));
            kind = (No % Np == 0) ? SO_PCAL_IMPLICIT : SO_PCAL_SHIFTING;
        }
    }
    p->kind = kind;
    p->fs_hz = fs;
    if (kind == SO_PCAL_TRIVIAL) {   /* :182-190, spacing arrives as int */
        int sp = (int)spacing;
        p->spacing_hz = sp;
        p->offset_hz = 0;
        p->nbins = (int)(fs / (double)pc_gcd(fabs((double)sp), fs));
        p->ntones = (int)floor(bw / sp);
    } else if (kind == SO_PCAL_SHIFTING) {   /* :311-330 */
        p->spacing_hz = spacing;
        p->offset_hz = offset;
        p->nbins = (int)(fs / (double)pc_gcd(fabs(spacing), fs));
        p->ntones = (int)floor((bw - offset) / spacing) + 1;
        p->rotatorlen = (long)(fs / (double)pc_gcd(fabs((double)offset), fs));
    } else {   /* :522-528 */
        p->spacing_hz = spacing;
        p->offset_hz = offset;
        p->nbins = (int)(fs / (double)pc_gcd(fabs((double)offset), fs));
        p->ntones = (int)floor((bw - offset) / spacing) + 1;
    }
    if (p->nbins <= 0 || p->ntones <= 0) return -1;
    p->buf = (double*)calloc((size_t)p->nbins * 4, sizeof(double));
    p->dft = (double*)calloc((size_t)p->nbins * 2, sizeof(double));
    if (kind == SO_PCAL_SHIFTING) {   /* :353-359 */
        double dphi = 2 * M_PI * (-(double)offset / fs);
        p->rot = (float*)malloc(sizeof(float) * 4 * (size_t)p->rotatorlen);
        for (long n = 0; n < 2 * p->rotatorlen; n++) {
            double arg = dphi * (double)n;
            p->rot[2 * n] = (float)cos(arg);
            p->rot[2 * n + 1] = (float)sin(arg);
        }
    }
    so_pcalx_clear(p);
    so_pcalx_adjust(p, sampleoffset);
    return 0;
}

void so_pcalx_free(so_pcalx* p)
{
    free(p->buf);
    free(p->dft);
    free(p->rot);
    memset(p, 0, sizeof(*p));
}

void so_pcalx_clear(so_pcalx* p)
{
    p->samplecount = 0;
    p->finalized = 0;
    memset(p->buf, 0, sizeof(double) * 4 * (size_t)p->nbins);
}

void so_pcalx_adjust(so_pcalx* p, size_t sampleoffset)
{
    p->pcal_index = sampleoffset % (size_t)p->nbins;
    p->rotator_index = (p->kind == SO_PCAL_SHIFTING) ? sampleoffset % (size_t)p->rotatorlen : 0;
}

/* dst[i] += x[i] (real) for i < n, into the complex buffer at index d */
static void add_real(double* buf, size_t d, const float* x, size_t n)
{
    for (size_t i = 0; i < n; i++) buf[2 * (d + i)] += (double)x[i];
}

/* dst[i] += x[i] * rot[i] for i < n */
static void add_rotated(double* buf, size_t d, const float* x, const float* rot, size_t n)
{
    for (size_t i = 0; i < n; i++) {
        buf[2 * (d + i)] += (double)x[i] * (double)rot[2 * i];
        buf[2 * (d + i) + 1] += (double)x[i] * (double)rot[2 * i + 1];
    }
}

int so_pcalx_extract(so_pcalx* p, const float* x, size_t len)
{
    if (p->finalized) return 0;
    const size_t N = (size_t)p->nbins;
    if (p->kind != SO_PCAL_SHIFTING) {   /* :270-291 and :609-637 */
        size_t tail = len % N, end = len - tail;
        const float* src = x;
        for (size_t n = 0; n < end; n += N, src += N) add_real(p->buf, p->pcal_index, src, N);
        if (tail) {
            add_real(p->buf, p->pcal_index, src, tail);
            p->pcal_index = (p->pcal_index + tail) % N;
        }
    } else {   /* :419-477: whole rotator periods add only whole pulses */
        const size_t L = (size_t)p->rotatorlen;
        size_t tail = len % L, end = len - tail;
        const float* src = x;
        const float* rot = p->rot + 2 * p->rotator_index;
        for (size_t n = 0; n < end; n += L, src += L)
            for (size_t q = 0; q < L / N; q++) add_rotated(p->buf, p->pcal_index, src + q * N, rot + 2 * q * N, N);
        if (tail) {
            size_t tail2 = tail % N, end2 = tail - tail2;
            for (size_t q = 0; q < end2 / N; q++) add_rotated(p->buf, p->pcal_index, src + q * N, rot + 2 * q * N, N);
            add_rotated(p->buf, p->pcal_index, src + end2, rot + 2 * end2, tail2);
            p->rotator_index = (p->rotator_index + tail) % L;
            p->pcal_index = (p->pcal_index + tail2) % N;
        }
    }
    p->samplecount += len;
    return 1;
}

uint64_t so_pcalx_final(so_pcalx* p, double* tones_out)
{
    const int N = p->nbins;
    if (!p->finalized) {
        p->finalized = 1;
        for (int b = 0; b < N; b++) {   /* fold the upper half, DFT with 1/N */
            p->buf[2 * b] += p->buf[2 * (N + b)];
            p->buf[2 * b + 1] += p->buf[2 * (N + b) + 1];
        }
    }
    size_t base = 0, step = 1;
    if (p->kind == SO_PCAL_SHIFTING) {   /* :503-512 */
        step = (size_t)floor(N * (p->spacing_hz / p->fs_hz));
    } else if (p->kind == SO_PCAL_IMPLICIT) {   /* :660-669: N*offset is an int product */
        step = (size_t)floor((double)N * (p->spacing_hz / p->fs_hz));
        int32_t prod = (int32_t)((uint32_t)N * (uint32_t)p->offset_hz);
        base = (size_t)(long long)floor((double)prod / p->fs_hz);
    }
    if (p->finalized == 1) {   /* the DFT (1/N) at the bins copied below */
        p->finalized = 2;
        for (int n = 0; n < p->ntones; n++) {
            size_t k = base + (size_t)n * step;
            if (k >= (size_t)N) break;
            double ar = 0, ai = 0;
            for (int b = 0; b < N; b++) {
                long long e = ((long long)k * b) % N;
                double a = -2.0 * M_PI * (double)e / (double)N;
                ar += p->buf[2 * b] * cos(a) - p->buf[2 * b + 1] * sin(a);
                ai += p->buf[2 * b] * sin(a) + p->buf[2 * b + 1] * cos(a);
            }
            p->dft[2 * k] = ar / N;
            p->dft[2 * k + 1] = ai / N;
        }
    }
    for (int n = 0; n < p->ntones; n++) {
        size_t idx = base + (size_t)n * step;
        if (idx >= (size_t)N) break;
        tones_out[2 * n] = p->dft[2 * idx];
        tones_out[2 * n + 1] = p->dft[2 * idx + 1];
    }
    return p->samplecount;
}

double so_pcal_delay64(const double* tones, int ntones, double spacing_hz, int upsample)
{
    const int G = upsample * ntones;
    double* mag = (double*)malloc(sizeof(double) * (size_t)G);
    int best = 0;
    for (int g = 0; g < G; g++) {
        double hr = 0, hi = 0;
        for (int n = 0; n < ntones; n++) {
            long long e = ((long long)n * g) % G;
            double a = 2.0 * M_PI * (double)e / (double)G;
            hr += tones[2 * n] * cos(a) - tones[2 * n + 1] * sin(a);
            hi += tones[2 * n] * sin(a) + tones[2 * n + 1] * cos(a);
        }
        mag[g] = sqrt(hr * hr + hi * hi);
        if (mag[g] > mag[best]) best = g;
    }
    double ym = mag[(best + G - 1) % G], y0 = mag[best], yp = mag[(best + 1) % G];
    double den = ym - 2 * y0 + yp;
    double d = (den != 0) ? 0.5 * (ym - yp) / den : 0.0;
    free(mag);
    double t = ((double)best + d) / (double)G;
    t -= floor(t);
    return t / spacing_hz;
}

/* ------------------------------------------------------------------------- */
/* doMaths (TaskCoreIPP.cpp:355-560)                                         */
/* ------------------------------------------------------------------------- */

int so_count_spectra(const so_settings* s, size_t remain, int* nffts)
{
    int acc = 0, spectra = 0, ffts = 0;
    while (remain >= s->raw_fullfft_bytes) {
        ffts++;
        remain -= s->raw_overlap_bytes;
        if (++acc >= s->core_overlapped_ffts) {
            spectra++;
            acc = 0;
            remain -= (s->raw_fullfft_bytes - s->raw_overlap_bytes);
        }
    }
    if (acc > 0) spectra++;
    if (nffts) *nffts = ffts;
    return spectra;
}

int so_domaths(const so_settings* s, int precision,
               const uint8_t* const* src_in, const size_t* len,
               void* const* out_auto_v, void* const* out_xpol_v, void* const* out_pcal_v)
{
    const long N = s->fft_points;
    const int ssb = s->fft_ssb_points;
    const int nsrc = s->num_sources;
    const int nx = s->num_xpols;
    const int K = s->core_overlapped_ffts;
    const int T = s->pcal_tonebins;
    const uint8_t* src[2];
    size_t remain = len[0];
    for (int r = 0; r < nsrc; r++) { src[r] = src_in[r]; if (len[r] < remain) remain = len[r]; }

    fplan* p = plan_new(N / 2, 1);
    if (!p) return -1;
    float* wtab = (float*)malloc(sizeof(float) * (size_t)N);
    double* wtab64 = (double*)malloc(sizeof(double) * (size_t)N);
    so_window_table(s->window, N, wtab);
    so_window_table64(s->window, N, wtab64);
    float* xs = (float*)malloc(sizeof(float) * (size_t)(N + 16));
    double* xd = (double*)malloc(sizeof(double) * (size_t)N);
    float* Xf[2]; double* Xd[2];
    for (int r = 0; r < 2; r++) {
        Xf[r] = (float*)malloc(sizeof(float) * 2 * (size_t)(ssb + 1));
        Xd[r] = (double*)malloc(sizeof(double) * 2 * (size_t)(ssb + 1));
    }
    float* zf = (float*)malloc(sizeof(float) * 2 * (size_t)N);
    double* zd = (double*)malloc(sizeof(double) * 2 * (size_t)N);

    /* output cursors; reset_spectrum (:275-291) zeroes the whole buffers */
    int maxsp = so_count_spectra(s, remain, NULL);
    if (maxsp < 1) maxsp = 1;
    size_t oa = 0, ox = 0, op = 0;  /* element offsets of the current spectrum */
    for (int r = 0; r < nsrc; r++) {
        if (precision == 32) memset(out_auto_v[r], 0, sizeof(float) * (size_t)ssb * maxsp);
        else memset(out_auto_v[r], 0, sizeof(double) * (size_t)ssb * maxsp);
        if (s->extract_pcal && out_pcal_v) {
            if (precision == 32) memset(out_pcal_v[r], 0, sizeof(float) * 2 * (size_t)T * maxsp);
            else memset(out_pcal_v[r], 0, sizeof(double) * 2 * (size_t)T * maxsp);
        }
    }
    for (int x = 0; x < nx; x++) {
        if (precision == 32) memset(out_xpol_v[x], 0, sizeof(float) * 2 * (size_t)ssb * maxsp);
        else memset(out_xpol_v[x], 0, sizeof(double) * 2 * (size_t)ssb * maxsp);
    }

    int curr_ffts = 0, acc = 0, spectra = 0;
    const float scale32 = (float)(1.0 / K);
    while (remain >= s->raw_fullfft_bytes) {
        for (int rs = 0; rs < nsrc; rs++) {
            int ch = rs == 0 ? s->use_channel_file1 : s->use_channel_file2;
            if (so_pick_unpacker(s) == 4)   /* the blanker may read past the FFT's bytes */
                so_vlba_unpack_mode(src[rs], (long)(len[rs] - (size_t)(src[rs] - src_in[rs])), (size_t)N, ch,
                                    s->vlba_fanout, s->bits_per_sample, xs);
            else if (so_pick_unpacker(s) == 5)
                so_mark5b_unpack(src[rs], (long)(len[rs] - (size_t)(src[rs] - src_in[rs])), (size_t)N, ch,
                                 s->source_channels, s->bits_per_sample, xs);
            else
                so_unpack(s, src[rs], xs, (size_t)N, ch);
            src[rs] += s->raw_overlap_bytes;
            int do_pcal = s->extract_pcal && out_pcal_v && ((curr_ffts % s->fft_overlap_factor) == 0);
            if (precision == 32) {
                if (do_pcal) so_extract_pcal32(s, xs, N, (float*)out_pcal_v[rs] + 2 * op);
                for (long i = 0; i < N; i++) xs[i] *= wtab[i];
                rfft32_plan(p, xs, Xf[rs], zf);
                float* oa32 = (float*)out_auto_v[rs] + oa;
                float re0 = Xf[rs][0], reN2 = Xf[rs][2 * (ssb - 1)];
                float acc0 = oa32[0] + re0 * re0;
                float accN = oa32[ssb - 1] + reN2 * reN2;
                for (int k = 1; k < ssb - 1; k++) {
                    float pr = Xf[rs][2 * k] * Xf[rs][2 * k] + Xf[rs][2 * k + 1] * Xf[rs][2 * k + 1];
                    oa32[k] = oa32[k] + pr;
                }
                oa32[0] = acc0;
                oa32[ssb - 1] = accN;
            } else {
                for (long i = 0; i < N; i++) xd[i] = (double)xs[i];
                if (do_pcal) so_extract_pcal64(s, xd, N, (double*)out_pcal_v[rs] + 2 * op);
                for (long i = 0; i < N; i++) xd[i] *= wtab64[i];
                rfft64_plan(p, xd, Xd[rs], zd);
                double* oa64 = (double*)out_auto_v[rs] + oa;
                for (int k = 0; k < ssb; k++)
                    oa64[k] += Xd[rs][2 * k] * Xd[rs][2 * k] + Xd[rs][2 * k + 1] * Xd[rs][2 * k + 1];
            }
            curr_ffts++;
        }
        remain -= s->raw_overlap_bytes;
        /* cross-spectra, pair (0,1) (:457-487) */
        for (int x = 0; x < nx; x++) {
            if (precision == 32) {
                float* o = (float*)out_xpol_v[x] + 2 * ox;
                float old0 = o[0], oldN = o[2 * (ssb - 1)];
                for (int k = 1; k < ssb - 1; k++) {
                    float ar = Xf[0][2 * k], ai = Xf[0][2 * k + 1];
                    float br = Xf[1][2 * k], bi = -Xf[1][2 * k + 1];
                    o[2 * k] = o[2 * k] + (ar * br - ai * bi);
                    o[2 * k + 1] = o[2 * k + 1] + (ar * bi + ai * br);
                }
                o[0] = old0 + Xf[0][0] * Xf[1][0];
                o[1] = 0;
                o[2 * (ssb - 1)] = oldN + Xf[0][2 * (ssb - 1)] * Xf[1][2 * (ssb - 1)];
                o[2 * (ssb - 1) + 1] = 0;
            } else {
                double* o = (double*)out_xpol_v[x] + 2 * ox;
                for (int k = 0; k < ssb; k++) {
                    double ar = Xd[0][2 * k], ai = Xd[0][2 * k + 1];
                    double br = Xd[1][2 * k], bi = -Xd[1][2 * k + 1];
                    o[2 * k] += ar * br - ai * bi;
                    o[2 * k + 1] += ar * bi + ai * br;
                }
            }
        }
        /* :489-511 */
        if (++acc >= K) {
            for (int rs = 0; rs < nsrc; rs++) {
                if (precision == 32) { float* o = (float*)out_auto_v[rs] + oa; for (int k = 0; k < ssb; k++) o[k] *= scale32; }
                else { double* o = (double*)out_auto_v[rs] + oa; for (int k = 0; k < ssb; k++) o[k] *= 1.0 / K; }
            }
            for (int x = 0; x < nx; x++) {
                if (precision == 32) { float* o = (float*)out_xpol_v[x] + 2 * ox; for (int k = 0; k < 2 * ssb; k++) o[k] *= scale32; }
                else { double* o = (double*)out_xpol_v[x] + 2 * ox; for (int k = 0; k < 2 * ssb; k++) o[k] *= 1.0 / K; }
            }
            oa += (size_t)ssb; ox += (size_t)ssb; op += (size_t)T;
            spectra++;
            acc = 0;
            for (int rs = 0; rs < nsrc; rs++) src[rs] += (s->raw_fullfft_bytes - s->raw_overlap_bytes);
            remain -= (s->raw_fullfft_bytes - s->raw_overlap_bytes);
        }
    }
    if (acc > 0) spectra++;   /* partial spectrum, unscaled (:519-524) */

    free(wtab); free(wtab64); free(xs); free(xd); free(zf); free(zd);
    for (int r = 0; r < 2; r++) { free(Xf[r]); free(Xd[r]); }
    plan_free(p);
    return spectra;
}

/* ------------------------------------------------------------------------- */
/* VLBA through mark5access 1.5.0 (vendored under mark5access/)              */
/* ------------------------------------------------------------------------- */

/* initmodulate (mark5access/mark5access/mark5_format_vlba.c:167-196): a
 * 16-stage shift register starting all ones; each step shifts in
 * b10 ^ b12 ^ b13 ^ b15 at b0 and that bit is the output; after every 8th
 * output one further step is taken unrecorded (the track parity bit). */
void so_vlba_modulate(uint8_t* out)
{
    unsigned reg = 0xFFFFu;
    for (int i = 0; i < SO_VLBA_PAYLOAD; i++) {
        for (int rep = 0; rep < ((i % 8 == 7) ? 2 : 1); rep++) {
            const unsigned k = ((reg >> 10) ^ (reg >> 12) ^ (reg >> 13) ^ (reg >> 15)) & 1u;
            reg = ((reg << 1) | k) & 0xFFFFu;
            if (rep == 0) out[i] = (uint8_t)k;
        }
    }
}

/* Mark5 fill pattern (blanker_mark5.c:41-46, little-endian host): the 64-bit
 * word 0x1122334411223344, i.e. bytes 44 33 22 11 repeated.  A word that is
 * not wholly inside the caller's `avail` bytes counts as data (the reference
 * reads past its buffer there). */
static int vl_fill(const uint8_t* p, long off, int bytes, long avail)
{
    static const uint8_t pat[4] = {0x44, 0x33, 0x22, 0x11};
    if (off < 0 || off + bytes > avail) return 0;
    for (int b = 0; b < bytes; b++)
        if (p[off + b] != pat[b & 3]) return 0;
    return 1;
}

/* blanker_mark5 (blanker_mark5.c:126-200) for blank zone z of a `payload`-
 * byte VLBA payload at byte `seg` of p.  The blanker cuts the payload's
 * 64-bit words into zones of 4096 words (32 KiB, log2blankzonesize 15):
 * 20000-byte payloads (8 tracks) have one zone, 40000 (16 tracks) two,
 * 80000 (32 tracks) three.  Bytes i with vs <= i < ve (payload-relative) are
 * data, the rest decode to 0.  Both boundary searches (findfirstinvalid
 * :48-81, findfirstvalid :83-116) are the same bisection over the zone's
 * 64-bit words plus a 32-bit straggler check. */
void so_vlba_blank_zone(const uint8_t* p, long avail, long seg, int payload, int z, int* vs, int* ve)
{
    const int nword = payload / 8;
    const int b = 4096 * z;
    const int e = (b + 4096 < nword) ? b + 4096 : nword;
    const int start_ok = !vl_fill(p, seg + 8L * b, 8, avail);
    const int end_ok = !vl_fill(p, seg + 8L * (e - 1), 8, avail);
    if (start_ok && end_ok) { *vs = 0; *ve = 1 << 30; return; }
    if (!start_ok && !end_ok) { *vs = 1 << 30; *ve = 0; return; }
    /* bisection for the boundary word: `want_fill` = the kind of word that
     * moves the boundary down */
    const int want_fill = start_ok;   /* data then fill: find first fill word */
    const int n = e - b;
    int i = n / 2, r = n - 1;
    for (int s = n / 4; s > 0; s >>= 1) {
        const int isfill = vl_fill(p, seg + 8L * (b + i), 8, avail);
        if (isfill == want_fill) { r = i; i -= s; }
        else i += s;
    }
    int byte = 8 * (b + r);
    if (start_ok) {
        if (vl_fill(p, seg + byte - 4, 4, avail)) byte -= 4;   /* data32[r/4 - 1] */
        *vs = 0;
        *ve = byte;
    } else {
        if (vl_fill(p, seg + byte - 8, 4, avail)) byte += 4;   /* data32[r/4 - 2] */
        *vs = byte;
        *ve = 1 << 30;
    }
}

/* One 20000-byte payload (8 tracks): its single zone. */
void so_vlba_blank(const uint8_t* p, long avail, long seg, int* vs, int* ve)
{
    so_vlba_blank_zone(p, avail, seg, SO_VLBA_PAYLOAD, 0, vs, ve);
}

/* mark5access's 32-track reordering of a little-endian 32-bit word
 * (mark5_format_vlba.c:74-76): bits of the 0xAA55AA55 mask stay, bits of
 * 0x55005500 move down 7 places, bits of 0x00AA00AA up 7. */
static uint32_t vl_reorder32(uint32_t x)
{
    return (x & 0xAA55AA55u) | ((x & 0x55005500u) >> 7) | ((x & 0x00AA00AAu) << 7);
}

/* VLBAUnpacker::extract_samples (src/IA-32/DataUnpackers.cpp:500-528) for the
 * VLBA modes with one byte per time sample, ntrack = 8 x fanout:
 *  - the request is cut into chunks of 2*(framebytes - payloadoffset) =
 *    2*(2520 - 12) * ntrack bytes (40128, 80256, 160512); each chunk is one
 *    mark5_unpack call on an unpacker stream without headers
 *    (mark5_stream_unpacker.c:112-129), so the payload index and the
 *    modulation restart at every chunk (and at every call, i.e. every FFT);
 *  - inside a chunk a payload is 2500 * ntrack bytes (databytes,
 *    mark5_format_vlba.c:9533); the blanker runs at the start of each
 *    payload (mark5_stream.c:147-196), the third payload of a chunk being its
 *    16 * ntrack-byte tail;
 *  - a frame word of `fanout` bytes carries `fanout` time samples of every
 *    channel; the word index m (payload byte / fanout) selects modulate[m],
 *    which negates the word's samples; a word whose first byte lies outside
 *    its zone's data range gives zeros;
 *  - levels (initluts :90-165): 2-bit {-HiMag, +1, -1, +HiMag}[sign bit +
 *    2 x magnitude bit], 1-bit {+1, -1}[bit].  Per mode, sample u of a word
 *    (u < fanout) of channel c reads:
 *      2-bit fanout 1 (vlba_decode_2bit_8track_fanout1_decimation1 :4537):
 *        byte 0, sign bit 2c, magnitude bit 2c+1;
 *      2-bit fanout 2 (_2bit_16track_fanout2 :5147, lut2bit2): byte c/2,
 *        sign bit 4(c%2)+u, magnitude bit that + 2;
 *      2-bit fanout 4 (_2bit_32track_fanout4 :5907): byte k of the word
 *        reordered by reorder32, k = 0, 2, 1, 3 for c = 0..3, sign bit 2u,
 *        magnitude bit 2u+1 (lut2bit1);
 *      1-bit fanout 1 (_1bit_8track_fanout1 :1362): byte 0, bit c;
 *      1-bit fanout 2 (_1bit_16track_fanout2 :2021): byte c/4, bit 2(c%4)+u;
 *      1-bit fanout 4 (_1bit_32track_fanout4 :2891): byte c%4, bit 2u + c/4. */
size_t so_vlba_unpack_mode(const uint8_t* src, long avail, size_t count, int channel, int fanout, int nbit,
                           float* dst)
{
    static uint8_t mod[SO_VLBA_PAYLOAD];
    static int mod_ready = 0;
    if (!mod_ready) { so_vlba_modulate(mod); mod_ready = 1; }
    if ((fanout != 1 && fanout != 2 && fanout != 4) || (nbit != 1 && nbit != 2)) return 0;
    const int nch = nbit == 2 ? 4 : 8;
    const int c = channel % nch;
    const int ntrack = 8 * fanout;
    const long payload = 2500L * ntrack;
    const long chunk = 2L * (2520 - 12) * ntrack;
    const float himag = 3.3359f;   /* OPTIMAL_2BIT_HIGH, mark5_stream.h:72 */
    const float lvl[4] = {-himag, 1.0f, -1.0f, himag};
    for (size_t c0 = 0; c0 < count; c0 += (size_t)chunk) {
        const size_t n = (count - c0 < (size_t)chunk) ? count - c0 : (size_t)chunk;
        int zs[3] = {0, 0, 0}, ze[3] = {0, 0, 0};
        for (size_t o = 0; o < n; o++) {
            const long i = (long)(o % (size_t)payload);     /* payload byte = time sample */
            if (i == 0) {
                for (int z = 0; z < (int)((payload / 8 + 4095) / 4096); z++)
                    so_vlba_blank_zone(src, avail, (long)(c0 + o), (int)payload, z, &zs[z], &ze[z]);
            }
            const long m = i / fanout;                      /* frame word */
            const long i0 = m * fanout;                     /* its first byte */
            const int u = (int)(i - i0);
            const int z = (int)(i0 >> 15);
            float v = 0.0f;
            if (i0 >= zs[z] && i0 < ze[z]) {
                const uint8_t* w = src + c0 + (o - (size_t)u);   /* the word's first byte */
                if (nbit == 2) {
                    int sb, byte;
                    if (fanout == 1) { byte = w[0]; sb = 2 * c; }
                    else if (fanout == 2) { byte = w[c / 2]; sb = 4 * (c % 2) + u; }
                    else {
                        const uint32_t x = (uint32_t)w[0] | ((uint32_t)w[1] << 8) | ((uint32_t)w[2] << 16) |
                                           ((uint32_t)w[3] << 24);
                        const int k = (c == 1) ? 2 : (c == 2) ? 1 : c;
                        byte = (int)((vl_reorder32(x) >> (8 * k)) & 0xFFu);
                        sb = 2 * u;
                    }
                    const int mb = (fanout == 2) ? sb + 2 : sb + 1;
                    v = lvl[((byte >> sb) & 1) + 2 * ((byte >> mb) & 1)];
                } else {
                    int bit, byte;
                    if (fanout == 1) { byte = w[0]; bit = c; }
                    else if (fanout == 2) { byte = w[c / 4]; bit = 2 * (c % 4) + u; }
                    else { byte = w[c % 4]; bit = 2 * u + c / 4; }
                    v = ((byte >> bit) & 1) ? -1.0f : 1.0f;
                }
                if (mod[m]) v = -v;
            }
            dst[c0 + o] = v;
        }
    }
    return count;
}

/* 8 tracks, fanout 1, 2-bit (VLBA1_1-<Mbps>-4-2). */
size_t so_vlba_unpack(const uint8_t* src, long avail, size_t count, int channel, float* dst)
{
    return so_vlba_unpack_mode(src, avail, count, channel, 1, 2, dst);
}

/* Mark5BUnpacker::extract_samples (DataUnpackers.cpp:707-751) over
 * mark5access's Mark5B decoders (mark5b_decode_<B>bitstream_<nbit>bit_
 * decimation1, mark5_format_mark5b.c:239-1911): see the header. */
size_t so_mark5b_unpack(const uint8_t* src, long avail, size_t count, int channel, int nchan, int nbit, float* dst)
{
    const int B = nchan * nbit;
    if ((nbit != 1 && nbit != 2) || (B != 2 && B != 4 && B != 8 && B != 16 && B != 32) || (nchan % 2) != 0) return 0;
    const int c = channel % nchan;
    const float himag = 3.3359f;
    const float lvl[4] = {-himag, 1.0f, -1.0f, himag};
    long cur = -1;
    int zs = 0, ze = 0;
    for (size_t t = 0; t < count; t++) {
        const unsigned long long p = (unsigned long long)t * (unsigned)B + (unsigned)(c * nbit);
        const long i0 = (long)(((unsigned long long)t * (unsigned)B) >> 3);   /* the sample's byte group */
        const long pl = i0 / SO_MARK5B_PAYLOAD;
        if (pl != cur) {
            so_vlba_blank_zone(src, avail, pl * SO_MARK5B_PAYLOAD, SO_MARK5B_PAYLOAD, 0, &zs, &ze);
            cur = pl;
        }
        const long ip = i0 - pl * SO_MARK5B_PAYLOAD;
        float v = 0.0f;
        if (ip >= zs && ip < ze) {
            const unsigned byte = src[p >> 3];
            const unsigned sh = (unsigned)(p & 7u);
            if (nbit == 2) v = lvl[((byte >> sh) & 1u) + 2u * ((byte >> (sh + 1u)) & 1u)];
            else v = ((byte >> sh) & 1u) ? -1.0f : 1.0f;
        }
        dst[t] = v;
    }
    for (size_t n = 0; n < 16 && n < count; n++) dst[n] = (n & 1) ? -himag : himag;
    return count;
}

/* ------------------------------------------------------------------------- */
/* Mk5B framing (FileSource.cpp:184-215)                                     */
/* ------------------------------------------------------------------------- */

long so_mk5b_strip(const uint8_t* frames, long nbytes, uint8_t* payload)
{
    const long fs = 10016, hl = 16;
    long out = 0;
    int bad = 0;
    for (long f = 0; f + fs <= nbytes; f += fs) {
        uint32_t sync = (uint32_t)frames[f] | ((uint32_t)frames[f + 1] << 8) |
                        ((uint32_t)frames[f + 2] << 16) | ((uint32_t)frames[f + 3] << 24);
        if (sync != 0xABADDEEDu) bad = 1;
        memcpy(payload + out, frames + f + hl, (size_t)(fs - hl));
        out += fs - hl;
    }
    return bad ? -1 : out;
}
