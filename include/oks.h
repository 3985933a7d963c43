/*
 * oks.h -- C ABI of the MI355X (gfx950) reconstruction backend for the
 * OpenKSpace `ReconStrategy` hot path.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns
 * an `int` status (OKS_OK == 0) and never lets a C++ exception cross the
 * boundary.  Citations are to the reference tree (Sigilweaver/OpenKSpace):
 *   recon/ = crates/openkspace-recon/src/,  io/ = crates/openkspace-io/src/.
 *
 * Data layout is the reference's (ndarray standard / C-contiguous):
 *   k-space  [nc, nz, ny, nx] complex64, interleaved {re, im}
 *            (io/ismrmrd/reader.rs:316, `Array4<Complex32>`)
 *   mask     [nz, ny, nx] uint8 0/1       (io/ismrmrd/reader.rs:334)
 *   image    [nz, ty, tx] float32         (recon/strategy.rs:26-34)
 *   maps     [nc, ny, nx] complex64       (recon/sensitivity.rs:37-41)
 */
#ifndef OKS_H
#define OKS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* num_complex::Complex32 is #[repr(C)] {re: f32, im: f32}. */
typedef struct oks_c32 {
    float re, im;
} oks_c32;

/* ---------------------------------------------------------------- status */
/* Codes mirror the reference error enums: IoError (io/error.rs:4-24),
 * GrappaError (recon/grappa.rs:540-547), SenseError (recon/sense.rs:38-48),
 * CsError (recon/cs.rs:43-51). */
enum oks_status {
    OKS_OK = 0,
    OKS_E_INCONSISTENT = 1,        /* IoError::Inconsistent */
    OKS_E_UNSUPPORTED = 2,         /* IoError::Unsupported  */
    OKS_E_MISSING_FIELD = 3,       /* IoError::MissingField */
    OKS_E_GRAPPA_BAD_CONFIG = 10,  /* GrappaError::BadConfig */
    OKS_E_GRAPPA_ACS_TOO_SMALL = 11, /* GrappaError::AcsTooSmall{need,got} */
    OKS_E_GRAPPA_CHOLESKY = 12,    /* GrappaError::CholeskyFailed */
    OKS_E_SENSE_BAD_ACCEL = 20,    /* SenseError::BadAcceleration */
    OKS_E_SENSE_SHAPE = 21,        /* SenseError::ShapeMismatch */
    OKS_E_SENSE_INDIVISIBLE = 22,  /* SenseError::IndivisibleNy */
    OKS_E_CS_SHAPE = 30,           /* CsError::ShapeMismatch */
    OKS_E_CS_ODD = 31,             /* CsError::OddDimension */
    OKS_E_INVALID_ARG = 40,        /* null pointer / impossible size */
    OKS_E_DEVICE = 50,             /* HIP runtime or launch failure */
    OKS_E_OOM = 51,                /* device allocation failed / admission refused */
    OKS_E_BUSY = 52                /* serving queue full */
};

/* Non-error fallbacks of the reference, reported as flags beside OKS_OK
 * (recon/prewhiten.rs:49-66, recon/oversampling.rs:51-60,
 *  recon/strategy.rs:163-187,349-355,389-395,602-647,829-842, recon/sense.rs:131-136). */
enum oks_flags {
    OKS_F_PREWHITEN_SKIPPED = 1u << 0,  /* from_noise_acqs returned None      */
    OKS_F_OS_SKIPPED = 1u << 1,         /* OversamplingRemover::new was None  */
    OKS_F_NO_PATTERN = 1u << 2,         /* detect_pattern None -> plain path  */
    OKS_F_GRAPPA_CALIB_FAILED = 1u << 3,/* >=1 slice left undersampled        */
    OKS_F_SENSE_FALLBACK = 1u << 4,     /* SENSE fell back to IFFT+RSS        */
    OKS_F_SENSE_PIXEL_FAILED = 1u << 5, /* >=1 pixel system not PD -> zero    */
    OKS_F_CS_FALLBACK = 1u << 6,        /* odd dims -> IFFT+RSS               */
    OKS_F_CROPPED = 1u << 7,            /* maybe_crop changed the shape       */
    OKS_F_HOMODYNE = 1u << 8            /* partial Fourier: homodyne path     */
};

/* FftMode (recon/strategy.rs:72-79). */
enum oks_fft_mode { OKS_FFT_AUTO = 0, OKS_FFT_2D = 1, OKS_FFT_3D = 2 };

/* SenseMapSource (recon/strategy.rs:469-474). */
enum oks_map_source { OKS_MAPS_WALSH = 0, OKS_MAPS_ESPIRIT = 1 };

/* --------------------------------------------------------- input "file" */
/* The subset of the ISMRMRD header the strategies read
 * (recon/strategy.rs:141-142,227-245; io/ismrmrd/reader.rs:187-202). */
typedef struct oks_header {
    int32_t encoded_x, encoded_y, encoded_z; /* encoding.encoded_matrix */
    int32_t recon_x, recon_y, recon_z;       /* encoding.recon_matrix   */
    int32_t is_3d_encoding;                  /* IsmrmrdFile::is_3d_encoding() */
} oks_header;

/* What `IsmrmrdFile::read_kspace_with_mask` (io/ismrmrd/reader.rs:225-468)
 * yields with an identity pre-process closure, plus the noise scans
 * `read_calibration` returns (io/ismrmrd/reader.rs:167-183).  Every sampled
 * readout occupies the full grid row (center_sample = nx/2, no discards), the
 * placement the reader performs for un-preprocessed full-length lines.
 * Pre-whitening and readout-oversampling removal are applied by the backend,
 * exactly where the reference's closure would apply them. */
typedef struct oks_kspace {
    const oks_c32* kspace;   /* [nc, nz, ny, nx]                              */
    const uint8_t* mask;     /* [nz, ny, nx], host memory (index metadata)    */
    int64_t nc, nz, ny, nx;
    const oks_c32* noise;    /* [nc, n_noise] concatenated noise samples / NULL */
    int64_t n_noise;
    oks_header hdr;
} oks_kspace;

/* Output `ImageVolume` (recon/strategy.rs:26-34). `data` (and `gfactor` when
 * requested) are caller-allocated with room for nz*ny*nx floats as returned by
 * oks_output_shape(). */
typedef struct oks_image {
    float* data;
    float* gfactor;          /* optional; SENSE only */
    int64_t nz, ny, nx;      /* filled by the call */
    uint32_t flags;          /* oks_flags */
} oks_image;

/* ------------------------------------------------- strategy descriptors */
/* Field-for-field mirrors of the reference strategy structs and their
 * Default impls. */
typedef struct oks_ifft_rss_desc {   /* IfftRss  recon/strategy.rs:86-106 */
    int32_t remove_oversampling, prewhiten, phase_correct, partial_fourier;
    int32_t fft_mode, crop_to_recon_matrix;
} oks_ifft_rss_desc;

typedef struct oks_grappa_desc {     /* GrappaRss recon/strategy.rs:275-299 */
    int32_t remove_oversampling, prewhiten, phase_correct;
    int32_t kernel_ky, kernel_kx;
    float ridge;
    int32_t fft_mode, crop_to_recon_matrix;
} oks_grappa_desc;

typedef struct oks_sense_desc {      /* SenseRss recon/strategy.rs:492-535 */
    int32_t remove_oversampling, prewhiten, phase_correct;
    int32_t map_source, walsh_window, walsh_iters;
    int32_t espirit_kernel;
    float espirit_threshold;
    int32_t espirit_iters;
    float ridge;
    int32_t compute_gfactor, fft_mode, crop_to_recon_matrix;
} oks_sense_desc;

typedef struct oks_cs_desc {         /* CsRss recon/strategy.rs:753-775 */
    int32_t remove_oversampling, prewhiten, phase_correct;
    int32_t iters;
    float lambda;
    int32_t fft_mode, crop_to_recon_matrix;
} oks_cs_desc;

/* Fill a descriptor with the reference `Default` values. */
void oks_ifft_rss_default(oks_ifft_rss_desc* d);
void oks_grappa_default(oks_grappa_desc* d);
void oks_sense_default(oks_sense_desc* d);
void oks_cs_default(oks_cs_desc* d);

/* Shape of the output volume: the grid's [nz, ny, nx] after
 * `IfftRss::maybe_crop` (recon/strategy.rs:227-253) when `crop` is set. */
void oks_output_shape(const oks_kspace* k, int32_t crop, int64_t out_shape[3]);

/* Sampling pattern (recon/grappa.rs:28-40). */
typedef struct oks_pattern {
    int64_t r, acs_start, acs_end, ky_lo, ky_hi;
} oks_pattern;

/* SamplingPattern::detect via detect_pattern (recon/grappa.rs:50-129,499-509):
 * OKS_OK and *out filled when a GRAPPA pattern is found, OKS_E_UNSUPPORTED
 * where the reference returns None.  Host-only integer logic. */
int oks_detect_pattern(const uint8_t* mask, int64_t nz, int64_t ny, int64_t nx, oks_pattern* out);

/* ----------------------------------------------------- readout placement */
/* An acquisition as read_kspace_with_mask and read_calibration read it: the
 * AcquisitionHeader fields they use (io/ismrmrd/acquisition.rs:15-64; slice
 * and segment are idx.slice and idx.segment) and where its samples start in
 * the data buffer ([active_channels][number_of_samples], channel-major,
 * complex64), in complex samples.  40 bytes. */
typedef struct oks_acq {
    uint64_t flags;  /* ISMRMRD flags (NON_IMAGE_MASK, ACQ_IS_REVERSE, ACQ_IS_PHASECORR_DATA) */
    uint16_t number_of_samples, active_channels, discard_pre, discard_post, center_sample;
    uint16_t kspace_encode_step_1, kspace_encode_step_2, slice, segment;
    uint16_t reserved[3];  /* zero */
    int64_t data_offset;
} oks_acq;

/* The encoding fields placement and oversampling removal read
 * (io/ismrmrd/header.rs: encoded_matrix, recon_matrix.x, ky/kz_limit.center). */
typedef struct oks_encoding {
    int32_t encoded_x, encoded_y, encoded_z, recon_x, ky_center, kz_center;
} oks_encoding;

/* Sizing pass of read_kspace_with_mask (io/ismrmrd/reader.rs:235-287): the
 * grid shape [nc, nz, ny, nx].  OKS_E_MISSING_FIELD where the reference
 * errors (no image acquisition, no channels).  Host arithmetic. */
int oks_place_shape(const oks_acq* acqs, int64_t n, const oks_encoding* enc, int64_t shape[4]);

/* ------------------------------------------------------------- context */
/* Opaque per-device state: stream, FFT plans, twiddles, workspaces, cached
 * HIP graphs (replaces the reference's per-call FftPlanner construction,
 * recon/fft.rs:65-66).  One owner at a time; calls on one ctx serialise. */
typedef struct oks_ctx oks_ctx;

int oks_ctx_create(int device, oks_ctx** out);
void oks_ctx_destroy(oks_ctx* ctx);
const char* oks_last_error(const oks_ctx* ctx);     /* message of the last failing call */
int oks_ctx_set_stream(oks_ctx* ctx, void* hip_stream); /* NULL = a private stream */
void* oks_ctx_stream(oks_ctx* ctx);
/* Options: "graphs" (0/1, capture the CS iteration loop), "zchunk" (slices per
 * IFFT+RSS chunk, >= 1), "time_stage" (an oks_stage, below), and the CS
 * column-pass launch shape "cs_sx" (columns per workgroup: 0 = automatic,
 * 4, 8, 16, 32) and "cs_threads" (128, 256, 512).  Launch shapes change
 * speed only, never results.  "rows_pipe" (0/1, default 1) selects the
 * pipelined IFFT+RSS row pass, which whitens after the kx IFFT instead of
 * before it (the same linear map, FP32 rounding in another order);
 * "cols_pipe" (0/1, default 1) lets column passes prefetch the next coil;
 * "rows_waves" (0 = automatic, or 8..16) sets the wavefronts per workgroup of
 * the one-row-per-workgroup row passes; "rows_wgs" (0 = as many as fit, 1,
 * 2) the persistent workgroups per CU of the pipelined row pass; "cols_sx"
 * (0 = automatic, or a power
 * of two up to 16) caps the columns per workgroup of the column passes and
 * "cols_threads" (0 = 512, or 256) sets their workgroup size (prefetching
 * needs 512).  Four switches (0/1, default 1) choose between two forms of
 * one computation whose FP32 rounding differs, both inside the strategies'
 * tolerances; 0 selects the earlier LDS-staged form: "cs_reg" (CS column
 * pass as register four-step transforms), "cols_reg" (the inverse column
 * passes likewise: the IFFT + RSS pass, and the complex-output passes at
 * 64 ... 512 points), "grappa_synth" (GRAPPA synthesis grouped by source row)
 * and "esp_rowt" (ESPIRiT per-voxel stage from the row-contracted Gram). */
int oks_ctx_set_option(oks_ctx* ctx, const char* name, int64_t value);
/* Launch shapes from the autotuner (python -m openkspace_amd.autotune, no
 * counterpart in the reference, whose FftPlanner plans per call,
 * recon/fft.rs:65-66): at context creation the library reads
 * $OKS_TUNE_FILE, or tuning/<arch>.json beside the library, a table of
 * (op, shape key) -> "cs_sx", "cs_threads", "cols_sx", "cols_threads",
 * "rows_waves".  Ops and keys: "ifft_rss" [nc, ny, nx, tx, nz if 3-D else 1]
 * for every IFFT + RSS tail, "cs" [ny, nx] for the FISTA passes.  None of
 * these options changes the computed bits; options set explicitly win, and
 * option "tuning" = 0 ignores the table.  oks_tune_lookup reads the table
 * for an architecture ("gfx950") without a device: OKS_OK and *value, or
 * OKS_E_MISSING_FIELD when the table, the entry or the option is absent. */
int oks_tune_lookup(const char* arch, const char* op, const int64_t* key, int32_t nkey, const char* name,
                    int64_t* value);
/* Current value of an option: the launch options above, "tuning",
 * "tuned_calls" (calls that took launch shapes from the table) and
 * "tuning_entries" (entries of this context's table, -1 without one). */
int oks_ctx_get_option(oks_ctx* ctx, const char* name, int64_t* value);
/* The device's architecture name without feature suffixes ("gfx950"). */
int oks_ctx_arch(oks_ctx* ctx, char* buf, int32_t cap);
int oks_ctx_synchronize(oks_ctx* ctx);
/* Flags raised on the device by the *_dev calls since the last take
 * (whitener not PD, SENSE pixel failure, GRAPPA calibration failure);
 * synchronises the stream and clears them. */
int oks_ctx_take_flags(oks_ctx* ctx, uint32_t* flags);

/* Stage timing (option "time_stage" = an oks_stage): every launch sequence of
 * that stage is bracketed by HIP events on the ctx stream and charged its
 * algorithmic work (bytes for the HBM-bound stages, flops for GRAPPA).
 * oks_ctx_stage_time synchronises and returns the totals since the option
 * was set. */
enum oks_stage {
    OKS_STAGE_NONE = 0,
    OKS_STAGE_IFFT_RSS = 1,          /* IFFT + crop + RSS tail; bytes: grid read + image write      */
    OKS_STAGE_GRAPPA_CALIBRATE = 2,  /* Gram, rhs, FP64 ridge solve; flops of recon/grappa.rs:156-308 */
    OKS_STAGE_GRAPPA_APPLY = 3,      /* synthesis GEMM; flops 8 x filled points x nc x n_src          */
    OKS_STAGE_CS_ITERATIONS = 4,     /* FISTA iterations; bytes 5*8 + 1 per pixel, plane, iteration   */
    OKS_STAGE_PLACE = 5,             /* readout placement; bytes: samples read + placed samples written */
    OKS_STAGE_ESPIRIT = 6            /* ESPIRiT maps; FP64 flops of the calibration Gram and the per-voxel
                                      * Gram and power iteration (the eigenpair search is not counted) */
};
int oks_ctx_stage_time(oks_ctx* ctx, double* ms, double* units, int64_t* launches);

/* Device memory and timing through the library's own HIP runtime. kind:
 * 0 host->device, 1 device->host, 2 device->device (synchronous). */
int oks_device_malloc(oks_ctx* ctx, size_t bytes, void** ptr);
int oks_device_free(oks_ctx* ctx, void* ptr);
int oks_memcpy(oks_ctx* ctx, void* dst, const void* src, size_t bytes, int32_t kind);
int oks_memset(oks_ctx* ctx, void* dst, int value, size_t bytes);
int oks_event_create(oks_ctx* ctx, void** ev);
int oks_event_record(oks_ctx* ctx, void* ev);          /* on the ctx stream */
int oks_event_elapsed_ms(oks_ctx* ctx, void* start, void* stop, float* ms);
int oks_event_destroy(oks_ctx* ctx, void* ev);

/* ---------------------------------------------------- strategy level */
/* One entry point per ReconStrategy impl: `reconstruct(&self, file)`
 * (recon/strategy.rs:58-67) with the file replaced by the tensors
 * read_kspace_with_mask and read_calibration produce.  The plain form takes
 * host buffers and blocks until the image is on the host; the _dev form takes
 * device kspace/noise/image (mask stays on the host), is asynchronous on the
 * ctx stream, and leaves device-raised flags for oks_ctx_take_flags.
 * Channel counts (the reference has no limit): ifft-rss, GRAPPA and CS take
 * up to 64 channels, SENSE up to 32; more fail with OKS_E_UNSUPPORTED. */
int oks_recon_ifft_rss(oks_ctx* ctx, const oks_ifft_rss_desc* d, const oks_kspace* k, oks_image* img);
int oks_recon_ifft_rss_dev(oks_ctx* ctx, const oks_ifft_rss_desc* d, const oks_kspace* k, oks_image* img);
    /* IfftRss::reconstruct, recon/strategy.rs:113-223 */
/* No counterpart in the reference: oks_recon_ifft_rss_dev for callers that
 * cannot allocate device memory (HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and flags; every
 * workspace comes from a static device array of OKS_NOALLOC_SCRATCH_BYTES per
 * device instead of the context's allocations, and a call whose workspaces do
 * not fit fails with OKS_E_UNSUPPORTED.  Calls that overlap on one device
 * share that array and race. */
#define OKS_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int oks_recon_ifft_rss_dev_noalloc(oks_ctx* ctx, const oks_ifft_rss_desc* d, const oks_kspace* k, oks_image* img);
int oks_recon_grappa(oks_ctx* ctx, const oks_grappa_desc* d, const oks_kspace* k, oks_image* img);
int oks_recon_grappa_dev(oks_ctx* ctx, const oks_grappa_desc* d, const oks_kspace* k, oks_image* img);
    /* GrappaRss::reconstruct, recon/strategy.rs:306-453 */
int oks_recon_sense(oks_ctx* ctx, const oks_sense_desc* d, const oks_kspace* k, oks_image* img);
int oks_recon_sense_dev(oks_ctx* ctx, const oks_sense_desc* d, const oks_kspace* k, oks_image* img);
    /* SenseRss::reconstruct, recon/strategy.rs:542-733 (Walsh or ESPIRiT maps) */
int oks_recon_cs(oks_ctx* ctx, const oks_cs_desc* d, const oks_kspace* k, oks_image* img);
int oks_recon_cs_dev(oks_ctx* ctx, const oks_cs_desc* d, const oks_kspace* k, oks_image* img);
    /* CsRss::reconstruct, recon/strategy.rs:782-891 */

/* ---------------------------------------------------- function level */
/* All device pointers, asynchronous on the ctx stream. */

/* ifft{1,2,3}_inplace / centred transforms (recon/fft.rs:19-56,
 * recon/cs.rs:141-218): centred transform of `data` (C-contiguous, `ndim`
 * dims) along `axes` in the order given; norm 0 none, 1 1/n (ifft_axis,
 * recon/fft.rs:73-83), 2 1/sqrt(n) (recon/cs.rs:166-176). */
int oks_fft_centered_dev(oks_ctx* ctx, oks_c32* data, const int64_t* shape, int32_t ndim, const int32_t* axes,
                         int32_t naxes, int32_t inverse, int32_t norm);
/* rss_combine_4d (recon/coil.rs:24-27): out[i] = sqrt(sum_c |in[c][i]|^2). */
int oks_rss_dev(oks_ctx* ctx, const oks_c32* in, int64_t nc, int64_t inner, float* out);
/* center_crop_3d (recon/crop.rs:13-27): each target clamped to [1, n], window
 * offset (n - t) / 2; in [nz, ny, nx] -> out [tz', ty', tx'], the clamped
 * shape written to out_shape. */
int oks_crop_dev(oks_ctx* ctx, const float* in, int64_t nz, int64_t ny, int64_t nx, int64_t tz, int64_t ty, int64_t tx,
                 float* out, int64_t out_shape[3]);
/* OversamplingRemover::apply (recon/oversampling.rs:103-154) on `rows`
 * readouts: in [rows, enc] -> out [rows, rec].  OKS_E_UNSUPPORTED where
 * OversamplingRemover::new returns None (recon/oversampling.rs:50-70: rec = 0,
 * enc <= rec, enc % rec != 0 or a ratio above 8). */
int oks_os_remove_dev(oks_ctx* ctx, const oks_c32* in, int64_t rows, int64_t enc, int64_t rec, oks_c32* out);
/* NoisePrewhitener::from_noise_acqs (recon/prewhiten.rs:48-111): W = chol(Psi)^-1
 * [nc, nc] lower, FP64 inside; W = I and a device flag when not PD.  nc <= 64
 * here and in oks_prewhiten_apply_dev (OKS_E_INVALID_ARG above). */
int oks_prewhiten_calibrate_dev(oks_ctx* ctx, const oks_c32* noise, int32_t nc, int64_t ns, oks_c32* W);
/* NoisePrewhitener::apply (recon/prewhiten.rs:118-142) on [nc, ns] in place. */
int oks_prewhiten_apply_dev(oks_ctx* ctx, const oks_c32* W, oks_c32* data, int32_t nc, int64_t ns);
/* GrappaKernel::calibrate (recon/grappa.rs:156-308): acs [nc, ny_acs, nx] ->
 * weights [r-1][nc][nc*kky*kkx]; OKS_E_GRAPPA_* as the reference errors. */
int oks_grappa_calibrate_dev(oks_ctx* ctx, const oks_c32* acs, int64_t nc, int64_t ny_acs, int64_t nx, int64_t r,
                             int64_t kky, int64_t kkx, float ridge, oks_c32* weights);
/* GrappaKernel::synthesize (recon/grappa.rs:318-424) on every kz slice of
 * kspace [nc, nz, ny, nx], in place, one weight set for all slices. */
int oks_grappa_synthesize_dev(oks_ctx* ctx, const oks_c32* weights, oks_c32* kspace, int64_t nc, int64_t nz,
                              int64_t ny, int64_t nx, int64_t r, int64_t kky, int64_t kkx);
/* sense_unfold_1d + sense_gfactor_1d (recon/sense.rs:64-162,181-243):
 * aliased, maps [nc, ny, nx] -> rho [ny, nx] (and g [ny, nx] when g != NULL);
 * R x R solves in FP64. */
int oks_sense_unfold_dev(oks_ctx* ctx, const oks_c32* aliased, const oks_c32* maps, int64_t nc, int64_t ny,
                         int64_t nx, int64_t r, float ridge, oks_c32* rho, float* g);
/* The per-kz loop of SenseRss over sense_unfold_1d + sense_gfactor_1d
 * (recon/strategy.rs:656-720) in one launch: aliased, maps [nc, nz, ny, nx]
 * -> rho [nz, ny, nx] (and g [nz, ny, nx] when g != NULL). */
int oks_sense_unfold_volume_dev(oks_ctx* ctx, const oks_c32* aliased, const oks_c32* maps, int64_t nc, int64_t nz,
                                int64_t ny, int64_t nx, int64_t r, float ridge, oks_c32* rho, float* g);
/* walsh_sensitivity_maps (recon/sensitivity.rs:37-142) per z: img, maps
 * [nc, nz, ny, nx]. */
int oks_walsh_maps_dev(oks_ctx* ctx, const oks_c32* img, int64_t nc, int64_t nz, int64_t ny, int64_t nx,
                       int64_t window, int64_t iters, oks_c32* maps);
/* espirit_sensitivity_maps (recon/espirit.rs:52-282) on one ACS block
 * acs [nc, acs_y, acs_x] -> maps [nc, ny, nx]; FP64 inside.  `kept` (may be
 * NULL) receives the number of calibration kernels above the threshold.
 * OKS_E_INVALID_ARG where the reference asserts (kernel larger than the ACS
 * block or the image), OKS_E_UNSUPPORTED beyond nc * kernel^2 = 2048 or a
 * kernel above 8. */
int oks_espirit_maps_dev(oks_ctx* ctx, const oks_c32* acs, int64_t nc, int64_t acs_y, int64_t acs_x, int64_t ny,
                         int64_t nx, int64_t kernel, float threshold, int64_t iters, oks_c32* maps, int64_t* kept);
/* read_kspace_with_mask (io/ismrmrd/reader.rs:225-468) with the strategies'
 * preprocess closure (recon/strategy.rs:153-161).  Every image acquisition is
 * whitened by W when it has W's channel count (W: device [w_nc, w_nc] or
 * NULL); with phase_correct, multiplied by the navigator correction of its
 * (slice, segment) that PhaseCorrector::from_phasecorr_acqs builds from the
 * table's ACQ_IS_PHASECORR_DATA acquisitions (recon/phasecorr.rs:61-169);
 * with remove_oversampling and a valid encoded_x / recon_x pair, it loses its
 * readout oversampling when it carries encoded_x samples; then it is placed:
 * DC offsets, centre sample and discards, first writer wins at the readout
 * centre, reversed readouts flipped.  acqs and mask are host memory, data
 * (n_data complex samples) and grid (the oks_place_shape shape) device
 * memory; mask [nz, ny, nx] and placed may be NULL.  OKS_E_INCONSISTENT when
 * an image acquisition's or navigator's samples fall outside data. */
int oks_place_acquisitions_dev(oks_ctx* ctx, const oks_acq* acqs, int64_t n, const oks_c32* data, int64_t n_data,
                               const oks_encoding* enc, const oks_c32* W, int32_t w_nc, int32_t phase_correct,
                               int32_t remove_oversampling, oks_c32* grid, uint8_t* mask, int64_t* placed);
/* fista_cs_single_coil (recon/cs.rs:59-139) on `planes` independent images:
 * kzf [planes, ny, nx], mask [mask_planes (1 or planes), ny, nx] -> out. */
int oks_fista_cs_dev(oks_ctx* ctx, const oks_c32* kzf, const uint8_t* mask, int64_t planes, int64_t mask_planes,
                     int64_t ny, int64_t nx, int64_t iters, float lambda, oks_c32* out);

/* -------------------------------------------------------- output writers */
enum oks_render_mode {
    OKS_RENDER_WINDOWED = 0, /* write_png_windowed (cli/main.rs:1112-1145)             */
    OKS_RENDER_LINEAR = 1    /* g-factor PNG: write_png_linear on [1, max(P99, 2)]
                                (cli/main.rs:866-870, 1095-1109)                     */
};
/* The 8-bit pixels of the CLI's PNG writers for every slice of vol [nz, ny, nx]
 * (device): per slice the window from percentile() (cli/main.rs:1147-1153,
 * exact order statistics of the slice) and byte = round(clamp((v - lo) /
 * (hi - lo), 0, 1)^0.9 * 255) (no exponent in the linear mode) -> out
 * [nz, ny, nx] u8 (device).  The PNG container itself stays with the host.
 * window (host, may be NULL) receives [nz][2] (lo, hi) and makes the call
 * synchronous.  OKS_E_INVALID_ARG for a window the CLI rejects
 * (cli/main.rs:625-628).  pct_low/pct_high are ignored in the linear mode. */
int oks_render_slices_dev(oks_ctx* ctx, const float* vol, int64_t nz, int64_t ny, int64_t nx, float pct_low,
                          float pct_high, int32_t mode, uint8_t* out, float* window);
/* write_nifti_volume's header (cli/main.rs:1033-1073): 348 header bytes and the
 * empty 4-byte extension block, 352 bytes into out; the volume follows as f32
 * little-endian, x fastest (the [nz, ny, nx] C order).  OKS_E_UNSUPPORTED
 * beyond i16 dimensions.  Host-only. */
int oks_nifti_header(int64_t nz, int64_t ny, int64_t nx, uint8_t* out);

/* ------------------------------------------------------------- serving */
/* No counterpart in the reference, whose CLI reconstructs one file per call
 * (cli/main.rs:678-784).  A request loop with continuous batching: requests
 * (one reconstruction each, k-space and noise resident in device memory) are
 * admitted by their working-set footprint against an HBM budget; arrived
 * requests with the same strategy, descriptor, shape, header and ky sampling
 * run together as one multi-slice reconstruction on one of `slots` streams,
 * while other slots run other batches.  The only asynchronous surface: it owns
 * its worker threads, contexts and streams. */
#define OKS_SERVER_MAX_SLOTS 16
#define OKS_SERVER_MAX_BATCH 64

enum oks_strategy_id {
    OKS_STRATEGY_IFFT_RSS = 0,
    OKS_STRATEGY_GRAPPA = 1,
    OKS_STRATEGY_SENSE = 2,
    OKS_STRATEGY_CS = 3
};

typedef struct oks_server oks_server;

typedef struct oks_server_config {
    int32_t device;
    int32_t slots;        /* batches in flight, one stream each (1..16)                */
    int32_t max_batch;    /* requests per batch (1..64); 1 = one request per launch chain */
    int32_t max_pending;  /* queued requests before submit returns OKS_E_BUSY; 0 = 65536 */
    int64_t hbm_budget;   /* bytes of batch working set in flight; 0 = half the free HBM */
    int32_t time_stage;   /* oks_stage timed on every slot (stats.stage_*), 0 = none    */
    int32_t no_graphs;    /* 1: launch iteration loops eagerly instead of as HIP graphs */
    int32_t reserved0;
    int64_t calib_cache_bytes; /* resident GRAPPA weights / SENSE maps shared by the slots,
                                  keyed by a hash of the calibration data; 0 = off (every
                                  request calibrates, as recon/strategy.rs:366-397 does) */
} oks_server_config;

typedef struct oks_request {
    uint64_t tag;         /* returned in the completion                                  */
    int32_t strategy;     /* oks_strategy_id                                             */
    const void* desc;     /* the strategy's descriptor; copied at submit                 */
    oks_kspace k;         /* kspace/noise in device memory, mask on the host (copied)     */
    float* out;           /* device image, oks_output_shape() floats                      */
    float* gfactor;       /* device g-factor map (SENSE) or NULL                          */
    double not_before;    /* arrival time on the server clock (s); earlier = now          */
} oks_request;

typedef struct oks_completion {
    uint64_t tag;
    int32_t status;       /* oks_status of the request's batch                          */
    uint32_t flags;       /* oks_flags of this request                                  */
    int64_t nz, ny, nx;   /* image shape                                                */
    int32_t batch, slot;  /* requests in its batch, slot it ran on                      */
    double t_arrival, t_start, t_done; /* server clock (s): latency = t_done - t_arrival */
} oks_completion;

typedef struct oks_server_stats {
    int64_t submitted, completed, failed, batches, pending;
    int64_t inflight_bytes, peak_inflight_bytes, hbm_budget, workspace_releases;
    int64_t slot_workspace_bytes[OKS_SERVER_MAX_SLOTS]; /* device workspace held per slot */
    double stage_ms, stage_units;     /* config.time_stage totals over all slots */
    int64_t stage_launches;
    int64_t calib_hits, calib_misses, calib_bytes; /* calibration cache: slices/volumes served
                                                     from it, calibrated afresh, bytes held */
} oks_server_stats;

void oks_server_default(oks_server_config* c);
int oks_server_create(const oks_server_config* cfg, oks_server** out);
void oks_server_destroy(oks_server* s);              /* finishes running batches, drops queued ones */
const char* oks_server_last_error(oks_server* s);
double oks_server_now(const oks_server* s);          /* server clock, seconds since create */
/* OKS_E_OOM when the request alone exceeds the budget, OKS_E_BUSY when full,
 * OKS_E_INVALID_ARG above 64 channels.  Requests of more than 32 channels
 * are never batched: each runs alone on the plain path. */
int oks_server_submit(oks_server* s, const oks_request* r);
/* Up to `max` completions; waits up to timeout_ms (<0 forever) for the first. */
int oks_server_poll(oks_server* s, oks_completion* out, int32_t max, int32_t timeout_ms, int32_t* n);
/* Wait until nothing is queued or running (timeout_ms < 0: forever); OKS_E_BUSY on timeout. */
int oks_server_drain(oks_server* s, int32_t timeout_ms);
int oks_server_get_stats(oks_server* s, oks_server_stats* out);
/* Working-set bytes the server charges a request at admission. */
int oks_request_footprint(const oks_request* r, int64_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* OKS_H */
