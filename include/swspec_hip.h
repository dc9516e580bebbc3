/*
 * swspec_hip.h -- C ABI of the MI355X spectrometer core (TaskCoreHIP).
 *
 * Drop-in boundary: each entry point replaces one virtual of the reference's
 * compute-core plugin interface, class TaskCore (src/TaskCore.h:36-89), whose
 * only complete implementation is TaskCoreIPP (src/IA-32/TaskCoreIPP.cpp).
 * The reference dispatcher (src/TaskDispatcher.cpp:92-247) drives a core with
 * prepare -> { run ; join ; combineResults | write } ... -> finalize from one
 * thread; the same call order applies here.  Plain C types only: pointers,
 * sizes, integers.  Return codes: 0 = ok, negative = error (text from
 * swspec_hip_last_error).
 */
#ifndef SWSPEC_HIP_H
#define SWSPEC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWSPEC_HIP_ABI_VERSION 2

/* error codes */
#define SWSPEC_OK 0
#define SWSPEC_EINVAL (-1)      /* bad argument or configuration */
#define SWSPEC_EUNSUPPORTED (-2)/* configuration the core does not implement */
#define SWSPEC_EDEVICE (-3)     /* HIP runtime error */
#define SWSPEC_ENOMEM (-4)      /* device allocation refused */
#define SWSPEC_ESTATE (-5)      /* call out of order (e.g. join without run) */

/* window / format enumerations keep the reference's values (src/Settings.h:45-47) */
enum swspec_window { SWSPEC_WIN_NONE = 0, SWSPEC_WIN_COSINE, SWSPEC_WIN_COSINE2, SWSPEC_WIN_HAMMING,
                     SWSPEC_WIN_HANN, SWSPEC_WIN_BLACKMAN };
enum swspec_format { SWSPEC_FMT_UNKNOWN = -1, SWSPEC_FMT_RAWSIGNED = 0, SWSPEC_FMT_RAWUNSIGNED,
                     SWSPEC_FMT_MK5B, SWSPEC_FMT_IBOB, SWSPEC_FMT_VDIF, SWSPEC_FMT_VLBA, SWSPEC_FMT_MKIV,
                     SWSPEC_FMT_MARK5B, SWSPEC_FMT_MAXIM };

/* flags */
#define SWSPEC_FLAG_NO_KAHAN 0x1u   /* plain FP32 accumulation (accuracy comparisons only) */
/* Mk5B input as stored on disk: whole 10016-byte frames (16-byte header +
 * 10000-byte payload).  The core skips the headers on the device instead of
 * FileSource::read stripping them on the host (src/FileSource.cpp:184-215);
 * every length passed to run()/run_device() is in frame bytes, and a
 * trailing partial frame is ignored.  Sync words are not checked. */
#define SWSPEC_FLAG_MK5B_FRAMED 0x2u
/* Open the column FFT with a radix-64 stage on the matrix cores: 64-point
 * DFTs as a split-FP16 product (x = hi + lo, W = hi + lo, the lo*lo term
 * dropped) accumulated in FP32 by v_mfma_f32_16x16x32_f16.  No counterpart
 * in the reference (IPP's DFT runs in FP32).  Needs an FFT length whose half
 * has a divisor that is a multiple of 64 with a cofactor <= 16384 (e.g.
 * 2^20 and 3.2M, not 32000); prepare fails otherwise.  Spectra stay within
 * the FP32 path's 1e-5 relative error of the FP64 result. */
#define SWSPEC_FLAG_MFMA64 0x4u

/*
 * Flat copy of the derived settings a core reads (src/Settings.h:59-159,
 * derived in src/swspectrometer.cpp:161-231).  All fields are the reference's
 * values after main() has run, e.g. use_channel_file1 is 0-based.
 */
typedef struct swspec_hip_cfg {
    uint32_t struct_size;          /* sizeof(swspec_hip_cfg) */
    uint32_t abi_version;          /* SWSPEC_HIP_ABI_VERSION */
    int64_t  fft_points;
    int32_t  fft_overlap_factor;
    int32_t  fft_ssb_points;
    int32_t  window;               /* swspec_window */
    int32_t  bits_per_sample;
    int32_t  source_channels;
    int32_t  channelorder_increasing;
    int32_t  source_format;        /* swspec_format */
    int32_t  use_channel_file1;    /* 0-based */
    int32_t  use_channel_file2;    /* 0-based */
    int32_t  num_sources;          /* 1 or 2 */
    int32_t  num_xpols;            /* 0 or 1 (pair 0,1) */
    int64_t  raw_fullfft_bytes;
    int64_t  raw_overlap_bytes;
    int32_t  core_overlapped_ffts; /* K: spectra are scaled by 1/K */
    int32_t  max_spectra_per_buffer;
    int32_t  extract_pcal;
    int32_t  pcal_tonebins;
    int32_t  pcal_rotatorlen;
    int32_t  device;               /* HIP device ordinal */
    double   samplingfreq;
    double   pcaloffsethz;
    double   pcalharmonicshz;
    int64_t  max_rawbuf_bytes;     /* largest raw buffer run() will be given */
    uint32_t flags;                /* SWSPEC_FLAG_* */
    uint32_t ffts_per_chunk;       /* 0 = automatic */
    char     mark5_format[32];     /* cfg->sourceformat_str for the mark5access formats
                                      (VLBA: "VLBA1_<fanout>-<Mbps>-<nchan>-<nbit>[/<decim>]",
                                      mark5_stream.c:556-571; decoded: fanout 1/2/4 with
                                      4 channels x 2 bits or 8 x 1 bit); NUL-terminated */
} swspec_hip_cfg;

typedef struct swspec_hip_core swspec_hip_core;

/*
 * INI-level settings -> derived settings, restating swspectrometer.cpp:67-231
 * (float arithmetic included).  max_source_buffer_mb < 0 means the key is
 * absent (platform default platform_max_raw_mb, src/Settings.h:30-36).
 * Returns 0, SWSPEC_EINVAL for an invalid channel (:148-155) or for
 * ExtractPCal with PCalOffsetHz == 0 (the reference divides by zero there).
 */
typedef struct swspec_hip_ini {
    int32_t num_cores;
    double  max_source_buffer_mb;
    int32_t platform_max_raw_mb;
    int64_t fft_points;
    float   fft_integ_seconds;
    int32_t fft_overlap_factor;
    int32_t window;
    float   bandwidth_hz;
    float   pcal_offset_hz;
    float   pcal_harmonics_hz;
    int32_t source_format;
    int32_t bits_per_sample;
    int32_t channelorder_increasing;
    int32_t source_channels;
    int32_t use_channel_file1;     /* 1-based, as in the INI */
    int32_t use_channel_file2;
    int32_t extract_pcal;
    int32_t num_sources;
    int32_t do_xpol;
} swspec_hip_ini;

typedef struct swspec_hip_derived {
    int32_t averaged_ffts;
    int32_t fft_overlap_points;
    int32_t fft_ssb_points;
    int64_t raw_fullfft_bytes;
    int64_t raw_overlap_bytes;
    int64_t rawbuf_size;
    int32_t max_spectra_per_buffer;
    int32_t max_specffts_per_buffer;
    int32_t max_buffers_per_spectrum;
    int32_t core_overlapped_ffts;
    int32_t core_averaged_ffts;
    int32_t extract_pcal;
    int32_t pcal_tonebins;
    int32_t pcal_rotatorlen;
    int32_t pcal_pulses;
    int32_t pcal_pulses_per_fft;
    int32_t num_xpols;
    int32_t use_channel_file1;     /* 0-based */
    int32_t use_channel_file2;
    double  samplingfreq;
    int64_t max_rawbuf_size;
} swspec_hip_derived;

int swspec_hip_derive(const swspec_hip_ini* ini, swspec_hip_derived* out);

/* Fill a cfg from ini + derived (convenience for hosts without the reference's main). */
int swspec_hip_cfg_from_ini(const swspec_hip_ini* ini, const swspec_hip_derived* d, swspec_hip_cfg* cfg);

/*
 * TaskCore::TaskCore(int rank) + TaskCore::prepare(settings)
 * (src/TaskCore.h:39,47; TaskCoreIPP.cpp:70-145): picks the decoder
 * (DataUnpackerFactory.h:38-108), builds window, FFT plan, phase-cal rotator
 * and device buffers.  Unlike the reference (which logs and continues with a
 * NULL unpacker, :79-83) an unsupported configuration fails here.
 */
int swspec_hip_create(const swspec_hip_cfg* cfg, int rank, swspec_hip_core** core);

/*
 * TaskCore::run (src/TaskCore.h:58; TaskCoreIPP.cpp:155-166), asynchronous.
 * in[s]      : host raw buffer of source s (in_len[s] bytes, Buffer::getLength)
 * out_auto[s]: host float array of max(spectra,1) * fft_ssb_points
 * out_xpol[0]: host float array of 2 * max(spectra,1) * fft_ssb_points (or NULL)
 * out_pcal[s]: host float array of 2 * max(spectra,1) * pcal_tonebins (or NULL)
 * Output arrays are written by the time swspec_hip_join returns; input
 * buffers must stay untouched until then (the dispatcher contract,
 * TaskDispatcher.cpp:119-139).  Pinned host memory (swspec_hip_host_alloc)
 * makes the copies asynchronous.
 */
int swspec_hip_run(swspec_hip_core* core, const void* const* in, const size_t* in_len,
                   float* const* out_auto, float* const* out_xpol, float* const* out_pcal);

/*
 * Several channels of source 0 in one pass: one job per entry of channels[]
 * (0-based), outputs out_auto[j] / out_pcal[j] per job, no cross-power.
 * No TaskCore counterpart: the reference runs one core per UseFile1Channel
 * value over the same file (src/swspectrometer.cpp:126, :162); this does the
 * same channels in one pass over the raw bytes.
 */
int swspec_hip_create_multichannel(const swspec_hip_cfg* cfg, int nchannels, const int* channels, int rank,
                                   swspec_hip_core** core);

/* Same as swspec_hip_run with device-resident inputs and outputs. */
int swspec_hip_run_device(swspec_hip_core* core, const void* const* d_in, const size_t* in_len,
                          float* const* d_out_auto, float* const* d_out_xpol, float* const* d_out_pcal);

/*
 * TaskCore::join (src/TaskCore.h:65; TaskCoreIPP.cpp:173-183): blocks until
 * the run completes.  *nspectra = complete + partial spectra (:519-524);
 * output lengths in bytes as Buffer::setLength would carry them (:527-541).
 * Returns 0 or a negative error (the reference's join() == -1).
 */
int swspec_hip_join(swspec_hip_core* core, int* nspectra, size_t* out_len_auto, size_t* out_len_xpol,
                    size_t* out_len_pcal);

/*
 * TaskCore::combineResults (src/TaskCore.h:81; TaskCoreIPP.cpp:311-349):
 * adds the core's last outputs (first spectrum of each) into the combined
 * host arrays and clears the core's outputs.  combined_out has num_sources
 * auto arrays followed by num_xpols cross arrays; lengths set as the
 * reference does (cross length = fft_ssb_points floats, :333).
 */
int swspec_hip_combine(swspec_hip_core* core, float* const* combined_out, float* const* combined_pcal,
                       size_t* combined_out_len, size_t* combined_pcal_len);

/* TaskCore::resetBuffer (src/TaskCore.h:87; TaskCoreIPP.cpp:297-300): zero floats. */
void swspec_hip_reset(void* buf, size_t bytes);

/*
 * TaskCore::combineResults (TaskCoreIPP.cpp:311-349) for device-resident
 * results: after a swspec_hip_run_device run has been joined, adds its first
 * spectrum of each output into the caller's DEVICE combined arrays (same
 * layout as swspec_hip_combine) and clears the core's device outputs, on the
 * core's stream; returns when done.  Lengths as swspec_hip_combine.  Keeps
 * combined spectra resident across buffers (no host round trip per buffer).
 */
int swspec_hip_combine_device(swspec_hip_core* core, float* const* d_combined_out, float* const* d_combined_pcal,
                              size_t* combined_out_len, size_t* combined_pcal_len);

/* TaskCore::resetBuffer (TaskCoreIPP.cpp:297-300) on device memory: zero the
 * floats of [d_buf, d_buf + bytes) with a fill kernel; returns when done. */
int swspec_hip_reset_device(void* d_buf, size_t bytes);

/* TaskCore::finalize (src/TaskCore.h:71; TaskCoreIPP.cpp:189-227): release all. */
int swspec_hip_finalize(swspec_hip_core* core);

/* Last error text of a core (or of the library when core == NULL). */
const char* swspec_hip_last_error(const swspec_hip_core* core);

/*
 * Phase-cal tones of spectrum `spectrum` of source `source` from the last run:
 * amplitude/phase source data in FP64, tone n = bin n*floor(T*spacing/fs) of
 * the 1/T-scaled forward DFT of the folded vector (PCal.cpp:489-515).
 * tones: 2*ntones doubles (re, im).  No counterpart call in TaskCore: the
 * reference writes only the folded vector; this exposes the FP64 extraction.
 */
int swspec_hip_pcal_tones(swspec_hip_core* core, int source, int spectrum, double* tones, int ntones);

/* Tones per source the core extracts (0 when phase-cal is off). */
int swspec_hip_pcal_ntones(const swspec_hip_core* core);

/* Internal calculation time of the last run in seconds (device events),
 * the figure TaskCoreIPP::finalize reports (:204-206). */
double swspec_hip_last_run_seconds(const swspec_hip_core* core);

/*
 * Per-kernel timing with HIP events recorded on the core's stream around each
 * launch (no TaskCore counterpart; the reference times whole doMaths calls,
 * TaskCoreIPP.cpp:390-391,516-517).  kernel: SWSPEC_K_* below.  Totals
 * accumulate over runs while profiling is on; set_profiling clears them.
 * on: 0 off, 1 every kernel, or SWSPEC_PROF_MASK | (1 << SWSPEC_K_...) | ...
 * for only the kernels named (each event pair costs the stream a little: the
 * 3.2M 4-channel step takes about 2% longer with every kernel timed).
 */
#define SWSPEC_PROF_MASK 0x100
#define SWSPEC_K_PASS1 0
#define SWSPEC_K_PASS2 1
#define SWSPEC_K_PCAL 2
#define SWSPEC_K_FINALIZE 3
int swspec_hip_set_profiling(swspec_hip_core* core, int on);
int swspec_hip_kernel_stats(const swspec_hip_core* core, int kernel, double* total_ms, int64_t* launches);

/*
 * ---- batching and serving --------------------------------------------------
 * No TaskCore counterpart: the reference runs one job (file set) per process.
 * A job is one channel of one device-resident raw stream, integrated over
 * `nffts` consecutive non-overlapped FFTs into one spectrum scaled by 1/nffts
 * (the reference's 1/K with K = nffts, TaskCoreIPP.cpp:124), with optional
 * phase-cal (folded vector + FP64 tones).  The server advances every active
 * job by one chunk of FFTs per step, grouping jobs of equal FFT length into
 * shared launches; partial integrations stay resident on the device between
 * steps; admission is capped by a device-memory budget and a job count.
 */
#define SWSPEC_EAGAIN (-6)      /* no capacity right now: retry after jobs finish */

typedef struct swspec_hip_job {
    uint32_t struct_size;          /* sizeof(swspec_hip_job) */
    int32_t  window;               /* swspec_window */
    int64_t  fft_points;
    int32_t  source_format;        /* RawSigned/RawUnsigned/Mk5B (payload) */
    int32_t  source_channels;
    int32_t  bits_per_sample;      /* 2 */
    int32_t  channelorder_increasing;
    int32_t  channel;              /* 0-based */
    int32_t  nffts;                /* FFTs integrated (>= 1) */
    int32_t  extract_pcal;
    int32_t  flags;                /* SWSPEC_FLAG_MFMA64 (other bits ignored) */
    double   samplingfreq;
    double   pcaloffsethz;
    double   pcalharmonicshz;
    const void* d_raw;             /* device bytes; FFT k starts at k * raw bytes per FFT */
    uint64_t raw_len;
    float*   d_out_auto;           /* fft_points/2 + 1 floats (device) */
    float*   d_out_pcal;           /* 2 * tone bins floats (device) or NULL */
    double*  d_tones;              /* 2 * tones doubles (device) or NULL */
} swspec_hip_job;

typedef struct swspec_hip_server swspec_hip_server;

/* hbm_budget_bytes <= 0: half of the free device memory.  max_active <= 0: 256. */
int swspec_hip_server_create(int device, int64_t hbm_budget_bytes, int max_active, swspec_hip_server** out);
/* Admit a job (allocates its resident state).  SWSPEC_EAGAIN when full. */
int swspec_hip_server_submit(swspec_hip_server* s, const swspec_hip_job* job, int64_t* job_id);
/* Enqueue one step: every active job advances by one chunk.  Non-blocking
 * (at most two steps in flight).  *nactive: jobs still active afterwards. */
int swspec_hip_server_step(swspec_hip_server* s, int* nactive);
/* Ids of jobs whose results are complete on the device (non-blocking). */
int swspec_hip_server_poll(swspec_hip_server* s, int64_t* done_ids, int max_ids, int* ndone);
/* Block until all enqueued work is complete. */
int swspec_hip_server_sync(swspec_hip_server* s);
/* Run all admitted jobs to completion.  use_graph: capture the whole batch
 * once into a hipGraph and replay it `replays` times (the job set is kept for
 * replays); otherwise step-by-step launches, once. */
int swspec_hip_server_run_batch(swspec_hip_server* s, int use_graph, int replays);
int swspec_hip_server_destroy(swspec_hip_server* s);
/* Counters: out[0..5] = steps enqueued, cross-stream waits (a job's batch
 * moved to the other stream and waited for its previous one), jobs admitted,
 * jobs completed, jobs active, device bytes of resident job state. */
int swspec_hip_server_stats(const swspec_hip_server* s, int64_t* out, int n);
const char* swspec_hip_server_last_error(const swspec_hip_server* s);

/* Pinned host memory for Buffer rings (hipHostMalloc) and its release. */
void* swspec_hip_host_alloc(size_t bytes);
void swspec_hip_host_free(void* p);

/* Decoder only (no counterpart call: DataUnpacker::extract_samples,
 * src/IA-32/DataUnpacker.h:32-38): unpack `count` samples of `channel` from
 * device raw bytes into device floats.  Used to check bit-exactness. */
int swspec_hip_unpack_device(const swspec_hip_cfg* cfg, const void* d_raw, int channel, int64_t count,
                             float* d_out);

/*
 * swspec_hip_unpack_device without device allocation, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * arguments and results as swspec_hip_unpack_device, but asynchronous on
 * `stream` (a hipStream_t; NULL = the null stream): the output is complete
 * once work queued after it on that stream has run.  VLBA decode keeps its
 * scratch (modulation bits, blank-zone table) in a fixed 1 MiB device array
 * per GPU, so a VLBA call decodes at most 1748938752 samples at fanout 1 or
 * 2 and 2331918336 at fanout 4 (130752 blank-zone entries) and fails with
 * SWSPEC_EINVAL beyond that; other formats use no scratch.  Calls that overlap on one GPU (two streams, two threads) share
 * that array and race; the first call on a device also copies the constant
 * modulation bits there (synchronously, once).
 */
int swspec_hip_unpack_device_noalloc(const swspec_hip_cfg* cfg, const void* d_raw, int channel, int64_t count,
                                     float* d_out, void* stream);

/* Plan geometry for an FFT length (no device work) under the plan-shaping
 * flags (SWSPEC_FLAG_MFMA64).  info[0..15] = N1, N2, pass-1 columns per
 * workgroup, pass-1 threads, pass-2 threads, row pairs, pass-1 / pass-2
 * points per thread, pass-1 / pass-2 composite-radix flags, pass-1 / pass-2
 * stages, MFMA radix-64 stage on, first pass-1 radix, single pass (short
 * FFTs: the whole M = N/2-point FFT in one kernel; then N1 = M, N2 = 1 and
 * the pass-1 fields describe that FFT), LDS padding of the penultimate row
 * stage's output blocks (0: none).  ninfo may be smaller: the first ninfo
 * entries are written. */
int swspec_hip_plan_info(int64_t fft_points, uint32_t flags, int32_t* info, int ninfo);

/* Autotuner (no counterpart in the reference, which has one FFT plan from
 * IPP).  Measures every distinct plan the planner's choices give for
 * fft_points -- column / row points per thread 16 or 32, composite radices
 * per pass, a cap on the columns per workgroup -- over nfft FFTs of seeded
 * single-channel 2-bit input on `device` (pass 1 + pass 2, best of 3),
 * rejects any whose spectrum differs from the planner's own plan by more
 * than 1e-5, and records the fastest: every later plan for that length
 * (cores, server jobs, plan_info) uses it.  knobs_out[4] = column points per
 * thread, column-count cap (0: none), composite mask, row points per thread;
 * times in ms for the chosen and for the planner's own plan; *ncand = plans
 * measured.  Allocates device memory while it runs. */
int swspec_hip_autotune(int device, int64_t fft_points, int32_t window, int32_t nfft, int32_t* knobs_out,
                        double* ms_best, double* ms_default, int32_t* ncand);
/* Forget every recorded choice (plans revert to the planner's own). */
void swspec_hip_autotune_clear(void);

/*
 * ---- phase-cal extractor classes --------------------------------------------
 * The reference's standalone extractor family (src/IA-32/PhaseCal/PCal.h:51-190,
 * PCal.cpp): PCal::getNew picks PCalExtractorTrivial (offset 0),
 * PCalExtractorImplicitShift (the offset's repeat length is a multiple of the
 * comb's) or PCalExtractorShifting (FP32 rotator).  Samples are float32 on the
 * device; accumulation is FP64 (the reference's is FP32); the Shifting class
 * keeps the reference's sample map, including the samples its whole-period
 * loop leaves out when rotatorlen is not a multiple of nbins (PCal.cpp:432-441).
 * The reference leaves the fold indices unset until adjustSampleOffset; here
 * create applies its sampleoffset as adjustSampleOffset does.
 */
typedef struct swspec_hip_pcal swspec_hip_pcal;
#define SWSPEC_PCAL_TRIVIAL 0
#define SWSPEC_PCAL_SHIFTING 1
#define SWSPEC_PCAL_IMPLICIT 2
/* PCal::getNew (PCal.cpp:91-110) when kind == -1, else the constructor of
 * one class (:182, :309, :520).  device: HIP device ordinal. */
int swspec_hip_pcal_create(double bandwidth_hz, double spacing_hz, int offset_hz, size_t sampleoffset, int kind,
                           int device, swspec_hip_pcal** out);
/* ~PCal */
void swspec_hip_pcal_destroy(swspec_hip_pcal* p);
/* getLength (tones), getNBins and the class picked; rotatorlen 0 unless Shifting. */
int swspec_hip_pcal_info(const swspec_hip_pcal* p, int* kind, int* ntones, int* nbins, int64_t* rotatorlen);
/* clear (PCal.h:86): zero the accumulation and the sample count. */
int swspec_hip_pcal_clear(swspec_hip_pcal* p);
/* adjustSampleOffset (PCal.h:95). */
int swspec_hip_pcal_adjust_offset(swspec_hip_pcal* p, size_t sampleoffset);
/* extractAndIntegrate (PCal.h:113): d_samples is device memory of len floats,
 * read asynchronously on `stream` (a hipStream_t; NULL = the null stream);
 * keep it unchanged until work queued after it on that stream has run.
 * SWSPEC_ESTATE once finalized (the reference returns false). */
int swspec_hip_pcal_extract(swspec_hip_pcal* p, const float* d_samples, size_t len, void* stream);
/* getFinalPCal (PCal.h:141): finalizes once, then writes the tone bins the
 * reference copies into out (complex float, host) and/or out64 (complex
 * double, host), either may be NULL; entries past the reference's bin-range
 * break are left as they were.  *samplecount = samples integrated. */
int swspec_hip_pcal_final(swspec_hip_pcal* p, float* out, double* out64, uint64_t* samplecount);
/* Tones getFinalPCal writes (<= getLength). */
int swspec_hip_pcal_copied_tones(const swspec_hip_pcal* p);
/* getSeconds (PCal.h:120): samples / fs. */
double swspec_hip_pcal_seconds(const swspec_hip_pcal* p);
/* Time-domain delay of the comb (the peak-position analysis PCal.h:15-20
 * describes, for which the reference has no code): peak of
 * |sum_n T_n e^{2 pi i n spacing t}| over one comb period on upsample*tones
 * points, refined by a parabola through its neighbours; seconds in
 * [0, 1/spacing).  Needs getFinalPCal first. */
int swspec_hip_pcal_delay(swspec_hip_pcal* p, int upsample, double* seconds);
/* Last error of an extractor (or of create when p == NULL). */
const char* swspec_hip_pcal_last_error(const swspec_hip_pcal* p);

#ifdef __cplusplus
}
#endif
#endif
