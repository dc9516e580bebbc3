/*
 * swspec_oracle.h -- CPU restatement of the swspec spectrometer hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP core:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it, and never as the thing that is measured or shipped.
 *
 * Every function names the reference file:line it restates (paths relative to
 * the reference repository root).  The FP64 functions are the numerical
 * oracle; the FP32 functions restate the reference's single-precision IPP
 * core so that its accumulation order and DC/Nyquist handling are reproduced.
 *
 * Parity status of each function (how it is pinned) is listed in its comment.
 */
#ifndef SWSPEC_ORACLE_H
#define SWSPEC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Enumerations keep the reference's numeric values (src/Settings.h:45-47). */
enum so_window { SO_WIN_NONE = 0, SO_WIN_COSINE, SO_WIN_COSINE2, SO_WIN_HAMMING, SO_WIN_HANN, SO_WIN_BLACKMAN };
enum so_format { SO_FMT_UNKNOWN = -1, SO_FMT_RAWSIGNED = 0, SO_FMT_RAWUNSIGNED, SO_FMT_MK5B, SO_FMT_IBOB,
                 SO_FMT_VDIF, SO_FMT_VLBA, SO_FMT_MKIV, SO_FMT_MARK5B, SO_FMT_MAXIM };

/* INI-level inputs (src/swspectrometer.cpp:67-145; defaults at :67-91). */
typedef struct so_ini {
    int    num_cores;
    double max_source_buffer_mb;   /* <0: key absent -> platform default */
    int    platform_max_raw_mb;    /* PLATFORM_MAX_RAW_BUF_SIZE_MB (src/Settings.h:30-36) */
    long   fft_points;
    float  fft_integ_seconds;
    int    fft_overlap_factor;
    int    window;                 /* so_window */
    float  bandwidth_hz;           /* fs = 2*BW (src/swspectrometer.cpp:115-117) */
    float  pcal_offset_hz;
    float  pcal_harmonics_hz;
    int    source_format;          /* so_format */
    int    bits_per_sample;
    int    channelorder_increasing;
    int    source_channels;
    int    use_channel_file1;      /* 1-based as in the INI */
    int    use_channel_file2;
    int    extract_pcal;
    int    num_sources;            /* 1 or 2 input files */
    int    do_xpol;
    int    vlba_fanout;            /* VLBA formats: <fanout> of "VLBA1_<fanout>-..." (1, 2, 4) */
} so_ini;

/* Derived settings (src/Settings.h:59-159, derivation src/swspectrometer.cpp:161-231). */
typedef struct so_settings {
    int    num_cores;
    long   fft_points;
    float  fft_integ_seconds;
    int    fft_overlap_factor;
    int    window;
    int    bits_per_sample;
    int    channelorder_increasing;
    int    source_channels;
    int    use_channel_file1;      /* 0-based after main()'s decrement (:162-163) */
    int    use_channel_file2;
    float  samplingfreq;
    float  pcaloffsethz;
    float  pcalharmonicshz;
    int    source_format;
    int    extract_pcal;
    int    num_sources;
    int    num_xpols;
    size_t max_rawbuf_size;
    float  dt, df;
    int    averaged_ffts;
    int    fft_overlap_points;
    int    fft_ssb_points;
    double rawbytes_per_channelsample;
    size_t raw_fullfft_bytes;
    size_t raw_overlap_bytes;
    size_t fft_bytes, fft_bytes_ssb, fft_bytes_xpol;
    size_t rawbuf_size;
    int    max_spectra_per_buffer;
    int    max_specffts_per_buffer;
    int    max_buffers_per_spectrum;
    int    core_overlapped_ffts;
    int    core_averaged_ffts;
    int    pcal_tonebins;
    int    pcal_rotatorlen;
    int    pcal_pulses;
    int    pcal_pulses_per_fft;
    size_t pcal_result_bytes;
    int    vlba_fanout;            /* copied from so_ini */
} so_settings;

/* Settings derivation, float arithmetic as in the reference
 * (src/swspectrometer.cpp:161-231, gcd src/Helpers.cpp:150-162).
 * Returns 0, or -1 for the reference's own exits (bad channel, :148-155)
 * and -2 for PCalOffsetHz == 0 with ExtractPCal (the reference divides by
 * zero in Helpers::gcd there).
 * Pinned: hand-computed tables for the five workload configurations. */
int so_derive(const so_ini* ini, so_settings* out);

/* Helpers::gcd (src/Helpers.cpp:150-162). b must be non-zero. */
long long so_gcd(long long a, long long b);

/* ---- decoders -------------------------------------------------------------
 * 2-bit value map (src/IA-32/DataUnpackers.cpp:212,274,334,362): field s ->
 * {+1, -1, +3.3359, -3.3359}[s].  Pinned: exhaustive closed-form tables. */
extern const float so_map_rev[4];

/* Which decoder the factory picks (src/IA-32/DataUnpackerFactory.h:38-108):
 * 1 = TwoBitUnpacker, 2 = TwoBitSinglechannelUnpacker, 3 = Mk5BUnpacker (4/8 ch),
 * 4 = VLBAUnpacker for the modes with one byte per time sample (4 channels x
 *     2 bits or 8 channels x 1 bit at fanout 1, 2 or 4; see so_vlba_unpack_mode),
 * 5 = Mark5BUnpacker (mark5access Mark5B, decimation 1; see so_mark5b_unpack),
 * 0 = none supported here (8/16-bit, MKIV/Mark5B via mark5access, other VLBA modes). */
int so_pick_unpacker(const so_settings* s);

/* Unpack `count` samples of 0-based `channel` starting at src.  Restates
 * TwoBitUnpacker (DataUnpackers.cpp:191-244), TwoBitSinglechannelUnpacker
 * (:261-305, its documented intent "oldest sample in the MSB", not the
 * out-of-bounds LUT of :277), Mk5BUnpacker 4/8-channel (:358-407).
 * Returns number of samples written (count rounded up to the unroll factor,
 * as the reference does), or 0 if no decoder applies. */
size_t so_unpack(const so_settings* s, const uint8_t* src, float* dst, size_t count, int channel);

/* ---- windows --------------------------------------------------------------
 * generate_windowfunction (src/IA-32/TaskCoreIPP.cpp:569-605).  None, Cosine,
 * Cosine2 follow the source exactly.  Hamming/Hann/Blackman(alpha=0.5) are
 * IPP-defined; Intel's published formulas are used (symmetric, N-1).
 * Parity for the IPP windows is pinned only to the published formula. */
void so_window_table(int type, long points, float* out);
/* FP64 version (same formulas, not rounded). */
void so_window_table64(int type, long points, double* out);

/* ---- FFT ------------------------------------------------------------------
 * Real forward DFT X[k] = sum_n x[n] e^{-2 pi i k n / N}, unnormalised (the
 * reference's ippsDFTFwd_RToPerm_32f with IPP_FFT_DIV_INV_BY_N,
 * TaskCoreIPP.cpp:116,430).  Output is N/2+1 complex bins (re,im interleaved).
 * N must be even with all prime factors in {2,3,5,7} or small.
 * Pinned: brute-force DFT at small N, numpy pocketfft at all sizes, Parseval. */
int so_rfft64(const double* x, long n, double* X);
int so_rfft32(const float* x, long n, float* X);
/* Complex forward DFT in place (interleaved), FP64. */
int so_cfft64(double* z, long n);

/* ---- the core's per-buffer maths ------------------------------------------
 * TaskCoreIPP::doMaths (src/IA-32/TaskCoreIPP.cpp:355-560) restated for one
 * call with num_sources raw buffers.  precision 32 restates the reference's
 * FP32 path (unpack -> pcal -> window -> FFT32 -> power -> accumulate, DC and
 * Nyquist handled as in :438-447 and :467-486, scale 1/K at :493-498);
 * precision 64 computes every stage in double (the FP64 oracle) and stores
 * double outputs.
 *   out_auto[s]: max_spectra * fft_ssb_points values (float or double)
 *   out_xpol[0]: max_spectra * fft_ssb_points complex (float or double pairs)
 *   out_pcal[s]: max_spectra * pcal_tonebins complex
 * Returns the number of spectra (full or partial, :519-524).  Outputs must hold
 * so_count_spectra(s, min(len)) spectra (at least one); they are zeroed first
 * as reset_spectrum (:275-291) does. */
int so_domaths(const so_settings* s, int precision,
               const uint8_t* const* src, const size_t* len,
               void* const* out_auto, void* const* out_xpol, void* const* out_pcal);

/* Loop arithmetic of doMaths (:394-524) without the maths: number of spectra
 * (full + partial) and FFTs per source that one buffer of `len` bytes gives. */
int so_count_spectra(const so_settings* s, size_t len, int* nffts);

/* extract_PCal (TaskCoreIPP.cpp:633-693): fold x*rotator into tonebins bins;
 * rotator e^{-2 pi i off n / fs} built in FP64 and rounded to FP32 (:129-136).
 * FP32 accumulate (reference) or FP64 accumulate (oracle) variants. */
void so_extract_pcal32(const so_settings* s, const float* x, long points, float* pcal);
void so_extract_pcal64(const so_settings* s, const double* x, long points, double* pcal);

/* Generic fold with a sample offset (PCalExtractorShifting::adjustSampleOffset
 * and extractAndIntegrate, PCal.cpp:402-479): sample n of the chunk uses
 * rotator phase (sampleoffset+n) mod rotatorlen and bin (sampleoffset+n) mod
 * nbins.  rot32 != 0 rounds the rotator to FP32 first (the core's table,
 * TaskCoreIPP.cpp:133-134); accumulation is FP64.  acc: nbins complex. */
void so_pcal_fold64(const double* x, long len, double fs_hz, double offset_hz, int nbins,
                    long rotatorlen, uint64_t sampleoffset, int rot32, double* acc);

/* Phase-cal tone extraction from a folded vector (PCal.cpp:489-515):
 * forward C2C DFT of length nbins scaled by 1/nbins, tone n taken from bin
 * n*floor(nbins*spacing/fs).  FP64.  tones_out: ntones complex.
 * Pinned: the reference PCal unit-test KAT (PCal.cpp:861-951, phases
 * -90 deg + 5 deg * tone within 0.1 deg) and a brute-force extraction. */
void so_pcal_tones64(const double* folded, int nbins, int ntones, double spacing_hz,
                     double fs_hz, double* tones_out);
/* Number of tones, PCalExtractorShifting (PCal.cpp:328). */
int so_pcal_ntones(double bandwidth_hz, double spacing_hz, double offset_hz);
/* FP64 direct reference (PCal.cpp:139-175): per-sample rotator phase and fold,
 * followed by so_pcal_tones64.  tones_out: ntones complex. */
void so_pcal_reference64(const double* x, long len, double bandwidth_hz, double spacing_hz,
                         double offset_hz, uint64_t sampleoffset, double* tones_out);

/* ---- The PCal extractor classes (src/IA-32/PhaseCal/PCal.cpp) ----
 * State machine of PCalExtractorTrivial (:182-302), PCalExtractorShifting
 * (:309-515) and PCalExtractorImplicitShift (:520-678), chosen by
 * PCal::getNew (:91-110) when kind < 0.  The loops follow the reference's
 * block structure (whole blocks at a fixed destination index into a buffer
 * of 2*nbins, the tail advancing the index, the upper half folded in at
 * getFinalPCal); the accumulation is FP64 (the reference's is FP32) and
 * the rotator is the reference's FP32-rounded table.  The reference leaves
 * pcal_index/rotator_index uninitialised until adjustSampleOffset; here the
 * constructor applies its sampleoffset argument as adjustSampleOffset would.
 * Pinned: the PCal.cpp unit test (:861-951) for each kind it can select, the
 * brute-force rotating fold (so_pcal_reference64) up to the documented
 * nbins scale, and impulse inputs for the Shifting drop rule. */
enum { SO_PCAL_TRIVIAL = 0, SO_PCAL_SHIFTING = 1, SO_PCAL_IMPLICIT = 2 };
typedef struct {
    int kind;
    double fs_hz, spacing_hz;
    int offset_hz;
    int nbins, ntones;
    long rotatorlen;
    size_t pcal_index, rotator_index;
    uint64_t samplecount;
    int finalized;
    double* buf;    /* 2*nbins complex accumulation buffer */
    float* rot;     /* 2*rotatorlen complex, FP32 (Shifting) */
    double* dft;    /* nbins complex, valid once finalized */
} so_pcalx;
/* kind: SO_PCAL_* or -1 for the factory.  Returns 0 or -1. */
int so_pcalx_init(so_pcalx* p, int kind, double bandwidth_hz, double spacing_hz, int offset_hz, size_t sampleoffset);
void so_pcalx_free(so_pcalx* p);
void so_pcalx_clear(so_pcalx* p);
void so_pcalx_adjust(so_pcalx* p, size_t sampleoffset);
/* 1 on success, 0 once finalized (extractAndIntegrate's bool). */
int so_pcalx_extract(so_pcalx* p, const float* x, size_t len);
/* getFinalPCal: writes the tone bins it selects into tones_out (ntones
 * complex; entries past the selection's break stay as they were) and
 * returns the sample count. */
uint64_t so_pcalx_final(so_pcalx* p, double* tones_out);
/* Time-domain delay from tone vectors (the peak-position analysis PCal.h:15-20
 * describes; the reference has no code for it): |h(t)|, h(t) = sum_n T_n
 * e^{+2 pi i n spacing t}, on G = upsample*ntones points over one comb period,
 * argmax (lowest index on ties), parabolic refinement of |h| through the
 * neighbours.  Returns t in [0, 1/spacing).  Parity unpinned against the
 * reference (no code there); pinned by delayed synthetic combs. */
double so_pcal_delay64(const double* tones, int ntones, double spacing_hz, int upsample);

/* ---- VLBA through mark5access (VLBAUnpacker, DataUnpackers.cpp:459-535) ----
 * Restated from the vendored mark5access 1.5.0 sources for the modes whose
 * time sample is one byte, the only ones VLBAUnpacker::extract_samples walks
 * consistently (it advances its source by the sample count, :515-522):
 * "VLBA1_<fanout>-<Mbps>-4-2" and "VLBA1_<fanout>-<Mbps>-8-1" at fanout 1, 2
 * and 4 (8, 16 and 32 tracks).  mark5access cannot be built here (its
 * sources include an autotools-generated config.h), so these functions are
 * pinned by closed forms and invariants only: the register's period and
 * balance, literal LUT bytes per mode, blank zones of synthetic fill patterns
 * against a linear scan and hand traces, and the per-chunk modulation
 * restart (tests/test_oracle_vlba.py). */
#define SO_VLBA_PAYLOAD 20000   /* data bytes per frame, 8 tracks */
#define SO_VLBA_CHUNK 40128     /* 2 * (framebytes 20160 - payloadoffset 96) */
/* modulate[0..19999] (mark5_format_vlba.c:167-196), values 0/1. */
void so_vlba_modulate(uint8_t* out);
/* Blank zone of the 20000-byte payload at byte `seg` of p (blanker_mark5.c:
 * 126-200): data where vs <= i < ve.  Words beyond `avail` count as data. */
void so_vlba_blank(const uint8_t* p, long avail, long seg, int* vs, int* ve);
/* extract_samples of `count` samples of `channel` (0..3) from src; `avail`
 * bytes are readable from src (the blanker looks up to 20000 bytes past the
 * last chunk's payload start).  8 tracks, fanout 1, 2-bit. */
size_t so_vlba_unpack(const uint8_t* src, long avail, size_t count, int channel, float* dst);
/* Blank zone z (words [4096 z, min(4096 z + 4096, payload/8))) of the
 * `payload`-byte payload at byte `seg` of p; bounds in payload bytes. */
void so_vlba_blank_zone(const uint8_t* p, long avail, long seg, int payload, int z, int* vs, int* ve);
/* extract_samples for any of the six one-byte modes: fanout 1/2/4, nbit 2
 * (4 channels) or 1 (8 channels); `count` a multiple of the fanout.
 * Returns count, or 0 for a mode outside the six. */
size_t so_vlba_unpack_mode(const uint8_t* src, long avail, size_t count, int channel, int fanout, int nbit,
                           float* dst);

/* ---- Mark5B through mark5access (Mark5BUnpacker, DataUnpackers.cpp:660-758) --
 * "Mark5B-<Mbps>-<nchan>-<nbit>", decimation 1, nchan even (canHandleConfig
 * :755-758), nbitstream B = nchan x nbit in {2, 4, 8, 16, 32}: one
 * mark5_unpack over the whole request (no chunking, :728-731) on an unpacker
 * stream without headers, 10000-byte payloads each with one blank zone
 * (blanker_mark5), no modulation.  Channel c of time sample t reads bit
 * t*B + c*nbit of the payload (LSB first; 2-bit: that bit is the sign, the
 * next the magnitude, {-HiMag, +1, -1, +HiMag}[sign + 2 mag], initluts
 * mark5_format_mark5b.c:59-100; 1-bit {+1, -1}[bit]); a sample whose byte
 * group (B/8 bytes, or the byte for B < 8) starts outside its zone's data
 * range is 0.  The reference then overwrites samples 0..15 of every request
 * with std::rand() signs x 3.3359 (:747-750), a stream no restatement can
 * follow: here they are the fixed pattern +3.3359, -3.3359, ... (even,
 * odd sample).  Returns count, or 0 for a configuration outside the above. */
#define SO_MARK5B_PAYLOAD 10000
size_t so_mark5b_unpack(const uint8_t* src, long avail, size_t count, int channel, int nchan, int nbit, float* dst);

/* ---- Mk5B framing -----------------------------------------------------------
 * FileSource::read framed mode (src/FileSource.cpp:184-215): strip the 16-byte
 * header of each 10016-byte frame.  Returns payload bytes written, or -1 if a
 * frame lacks the 0xABADDEED sync word (the reference only warns, :401-427;
 * here it is reported). */
long so_mk5b_strip(const uint8_t* frames, long nbytes, uint8_t* payload);

#ifdef __cplusplus
}
#endif
#endif
