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
 *  recon/strategy.rs:349-355,389-395,602-647,829-842, recon/sense.rs:131-136). */
enum oks_flags {
    OKS_F_PREWHITEN_SKIPPED = 1u << 0,  /* from_noise_acqs returned None      */
    OKS_F_OS_SKIPPED = 1u << 1,         /* OversamplingRemover::new was None  */
    OKS_F_NO_PATTERN = 1u << 2,         /* detect_pattern None -> plain path  */
    OKS_F_GRAPPA_CALIB_FAILED = 1u << 3,/* >=1 slice left undersampled        */
    OKS_F_SENSE_FALLBACK = 1u << 4,     /* SENSE fell back to IFFT+RSS        */
    OKS_F_SENSE_PIXEL_FAILED = 1u << 5, /* >=1 pixel system not PD -> zero    */
    OKS_F_CS_FALLBACK = 1u << 6,        /* odd dims -> IFFT+RSS               */
    OKS_F_CROPPED = 1u << 7             /* maybe_crop changed the shape       */
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

#ifdef __cplusplus
}
#endif

#endif /* OKS_H */
