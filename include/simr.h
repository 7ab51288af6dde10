/*
 * simr.h - stage-level C ABI of the MI355X 3D-SIM reconstruction engine.
 *
 * This is the seam between host orchestration and device code.  It replaces
 * the reference's internal stage API (src/cudaSirecon/gpuFunctions.h:19-80 and
 * the two cuFFT call sites in src/cudaSirecon/cudaSirecon.cpp:238-257,542-579),
 * which took std::vector<GPUBuffer>* plus a global ReconParams and kept state in
 * __constant__ symbols.  Here every call takes an explicit context that owns its
 * device arena, HIP stream and FFT plans; no process-global state is used, so
 * any number of contexts can live in one process.
 *
 * All functions return 0 (SIMR_OK) on success and a negative simr_status on
 * failure; simr_last_error() returns a thread-local message.  Nothing throws
 * across this boundary.
 */
#ifndef SIMR_H
#define SIMR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SIMR_MAX_DIRS 8
#define SIMR_MAX_ORDERS 3 /* norders <= 3, i.e. nphases <= 5, as in the reference GPU path */
#define SIMR_MAX_PHASES 5

typedef enum simr_status {
  SIMR_OK = 0,
  SIMR_ERR_ARG = -1,       /* invalid argument / unsupported parameter combination */
  SIMR_ERR_HIP = -2,       /* HIP runtime error */
  SIMR_ERR_NOMEM = -3,     /* device arena exhausted / allocation failure */
  SIMR_ERR_IO = -4,        /* file could not be read or written */
  SIMR_ERR_FORMAT = -5,    /* malformed MRC/TIFF/config content */
  SIMR_ERR_STATE = -6,     /* call out of order (e.g. result requested before processing) */
  SIMR_ERR_UNSUPPORTED = -7
} simr_status;

/*
 * Reconstruction parameters.  Field names and meanings follow the reference's
 * ReconParams / ImageParams (src/cudaSirecon/cudaSireconImpl.h:104-195).
 * simr_params_default() fills the values the cudasirecon CLI uses when an
 * option is not given (src/cudaSirecon/cudaSirecon.cpp:16-72 overlaid by the
 * option defaults at 918-1005).
 */
typedef struct simr_params {
  /* image geometry */
  int32_t nx, ny, nz;          /* nz = z sections per (direction, phase) */
  int32_t ndirs, nphases, norders_output;
  float dxy, dz;               /* pixel sizes in microns */
  int32_t wave;                /* emission wavelength in nm */
  int32_t ntimes;              /* number of time points in the series */
  /* illumination pattern */
  float k0startangle, linespacing, na, nimm;
  int32_t nk0angles;
  float k0angles[SIMR_MAX_DIRS];
  /* algorithm */
  float otfcutoff, zoomfact;
  int32_t z_zoom;
  float explodefact;
  int32_t bFilteroverlaps, recalcarrays, napodize, bSearchforvector, bUseTime0k0;
  int32_t apodizeoutput;
  float apoGamma;
  int32_t bSuppress_singularities, suppression_radius, bDampenOrder0, bFitallphases;
  int32_t do_rescale, equalizez, equalizet, bNoKz0;
  float wiener;
  int32_t nforceamp;
  float forceamp[SIMR_MAX_ORDERS];
  /* OTF geometry.  bRadAvgOTF: rotationally averaged, nyotf == 1, table
   * [kr][kz] (nxotf radial x nzotf axial samples).  Otherwise the full OTF
   * over kx >= 0, table [kz][ky][kx] (nxotf x nyotf x nzotf, ky and kz in FFT
   * order, dkrotf the kx and ky step), read by trilinear interpolation with
   * the Hermitian mirror for kx < 0 (the reference's GPU path loads such
   * tables but evaluates them as zero, gpuFunctionsImpl.cu:1680-1714). */
  int32_t bRadAvgOTF, bOneOTFperAngle;
  int32_t nxotf, nyotf, nzotf;
  float dkrotf, dkzotf;
  /* camera */
  float constbkgd;
  int32_t bFastSIM;
  /* engine options (no counterpart in the reference) */
  int32_t fit_fp64;            /* 1: k0/modamp fit reductions in FP64 (default); 0: FP32 partials */
  /* lateral resampling of the raw sections before the reconstruction
     (deskewOneSection, cudaSirecon.cpp:1733-1760): raw sections are nx_raw
     wide; output column x of z section z samples the flat-fielded raw row at
     x - nx/2 + deskew_shift - deskew_factor (z - nz/2) + nx_raw/2 (linear
     interpolation), or at x when deskew_factor is 0 (a crop keeps the left
     nx columns).  nx_raw 0 or nx with deskew_factor 0: no resampling. */
  int32_t nx_raw;              /* width of the raw sections (0: nx) */
  float deskew_factor;         /* cos(deskew angle) * raw dz / dxy (0: none) */
  int32_t deskew_shift;        /* --deskewshift */
  /* Bessel-sheet SIM (--bessel, --besselNA, --besselExWave): k0 fitted on
     order 1 in 3D, order-0 weight 4, overlap OTF cutoff 0.01, axial cutoffs
     from the excitation's kz extent 2 NA / lambda, anisotropic apodization
     (gpuFunctionsImpl.cu:317-323, 505-506, 533-539, 554-555, 1260-1271, 1294,
     1570-1577) */
  int32_t bBessel;
  float BesselNA;              /* excitation NA of the Bessel beam */
  float BesselLambdaEx;        /* excitation wavelength in microns */
  /* 1 (default): a time point that takes time 0's k0 (ntimes > 1, t > 0,
     bUseTime0k0) runs its modamp fit, filter and assembly as one captured
     HIP graph per context, with the modamps computed on the device; 0: the
     same kernels launched one by one, the modamps computed on the host.
     Both give the same result bit for bit.  (No counterpart in the reference.) */
  int32_t use_graphs;
  int32_t reserved[8];
} simr_params;

/* Per-direction fit results, the observable outputs the reference prints. */
typedef struct simr_fit {
  float k0x, k0y;              /* modulation vector of order 1, in 1/micron */
  float k0guess_x, k0guess_y;  /* initial guess, 1/micron */
  float amp_re[SIMR_MAX_ORDERS], amp_im[SIMR_MAX_ORDERS];
  float corr[SIMR_MAX_ORDERS]; /* correlation coefficient of each order's modamp fit */
  int32_t n_angle_evals, n_mag_evals, n_reductions;
} simr_fit;

typedef struct simr_ctx simr_ctx;

void simr_params_default(simr_params* p);
const char* simr_last_error(void);
void simr_set_last_error(const char* msg);
const char* simr_version(void);
/* sizeof(simr_params) and sizeof(simr_fit), for binding-layout checks. */
size_t simr_params_size(void);
size_t simr_fit_size(void);

/* Number of complex OTF entries per (direction, order) for these parameters. */
size_t simr_otf_size(const simr_params* p);
/* Complex entries of the whole OTF argument of simr_create: (ndirs if bOneOTFperAngle else 1) x norders x simr_otf_size. */
size_t simr_otf_total(const simr_params* p);
/* Device bytes one context needs for these parameters (arena footprint). */
size_t simr_footprint(const simr_params* p);

/*
 * Create a reconstruction context.  otf: interleaved complex float, laid out
 * [nDirsOTF][norders][simr_otf_size()] with nDirsOTF = bOneOTFperAngle ? ndirs : 1.
 * device: HIP device ordinal.  The context owns a stream and a device arena.
 */
int simr_create(const simr_params* p, const float* otf, int device, simr_ctx** out);
int simr_destroy(simr_ctx* ctx);
int simr_get_params(const simr_ctx* ctx, simr_params* out);
/* Replace the parameters of a live context (wiener, k0 guesses, flags, ntimes,
 * ...).  Sizes that lay out device memory (image and zoomed dimensions, ndirs,
 * nphases, the OTF table geometry) must not change: SIMR_ERR_ARG otherwise.
 * Fit state of a time series (the time-0 k0) is kept. */
int simr_update_params(simr_ctx* ctx, const simr_params* p);

/*
 * Raw input: nImages = nz * ndirs * nphases sections of nx*ny pixels, in the
 * reference's section order (Angle->Z->Phase, or Z->Angle->Phase with bFastSIM).
 * The *_host variants copy from host memory; the *_dev variants take a device
 * pointer (already resident in HBM) and copy device-to-device on the context stream.
 */
int simr_set_raw_f32_host(simr_ctx* ctx, const float* raw);
int simr_set_raw_u16_host(simr_ctx* ctx, const uint16_t* raw);
int simr_set_raw_f32_dev(simr_ctx* ctx, const float* raw_dev);
int simr_set_raw_u16_dev(simr_ctx* ctx, const uint16_t* raw_dev);
/* Optional flat-field maps (background, slope), nx_raw*ny floats each (nx*ny without
 * resampling); NULL resets to constbkgd/1. */
int simr_set_flatfield(simr_ctx* ctx, const float* background, const float* slope);

/* Stage calls (in this order per time point). */
int simr_load_and_rescale(simr_ctx* ctx, int it);   /* flatfield + bleach correction + apodize + XY FFT + separation + z FFT */
int simr_set_time_index(simr_ctx* ctx, int it);
int simr_find_modulation(simr_ctx* ctx);             /* k0 search / fit and modamps for all directions */
int simr_filter_assemble(simr_ctx* ctx);             /* Wiener filtering + zoomed assembly into the output volume */
int simr_process_volume(simr_ctx* ctx);              /* find_modulation + filter_assemble */

/*
 * Page-locked host memory for stacks and results (no counterpart in the
 * reference, whose loaders read into pageable buffers): the *_host calls
 * above and below copy from and to such buffers by DMA at the link's rate,
 * where pageable memory goes through the runtime's staging copies.  NULL on
 * failure (simr_last_error tells why).
 */
void* simr_host_alloc(size_t bytes);
int simr_host_free(void* p);

/* Results. out: (zoom*nx)*(zoom*ny)*(z_zoom*nz) floats, layout [z][y][x]. */
size_t simr_result_elems(const simr_ctx* ctx);
int simr_get_result_host(simr_ctx* ctx, float* out);
/* Copies the result into caller-owned device memory (on the context stream, then synchronises). */
int simr_get_result_dev(simr_ctx* ctx, float* out_dev);
const float* simr_result_dev(simr_ctx* ctx);
int simr_get_fit(const simr_ctx* ctx, int dir, simr_fit* out);
/* Override k0 (1/micron) and modamps of one direction (known-k0 processing, parity checks). */
int simr_set_fit(simr_ctx* ctx, int dir, const simr_fit* in);

/*
 * Time-0 state of a time series: the k0 of every direction fitted at t = 0
 * (which t > 0 reuses when bUseTime0k0, cudaSirecon.cpp:296-331) and the
 * bleach-correction reference sums of t = 0 (which t > 0 reuses when
 * equalizet, gpuFunctionsImpl.cu:140-148).  Copying it from the context that
 * processed t = 0 into other contexts of the same shape lets later time
 * points run on any of them with the results of one sequential context.
 */
size_t simr_time0_state_size(const simr_ctx* ctx);
int simr_get_time0_state(simr_ctx* ctx, void* host, size_t bytes);
int simr_set_time0_state(simr_ctx* ctx, const void* host, size_t bytes);

/* Stage inspection (parity tests): copy one separated band of one direction
 * (half-complex spectrum [z][y][nx/2+1], complex float) to the host. */
int simr_get_band_host(simr_ctx* ctx, int dir, int band, float* out);
int simr_set_band_host(simr_ctx* ctx, int dir, int band, const float* in);
/* Run only the Wiener filter of one direction, in place on its bands. */
int simr_filter_dir(simr_ctx* ctx, int dir);
/* Build the overlap pair of (order1, order2) at k0 (1/micron) from the current
 * bands and return its z-summed products: P = sum_z conj(X) Y (complex double
 * [ny][nx]), SX = sum_z |X|^2, SY = sum_z |Y|^2 (double [ny][nx]). */
int simr_debug_overlaps(simr_ctx* ctx, int dir, int order1, int order2, float k0x, float k0y, double* P,
                        double* sx, double* sy);
/* The same overlap pair in real space, for the --saveoverlaps dump
 * (cudaSirecon.cpp:330-352,433-445; makeoverlaps' inverse 3D FFT,
 * gpuFunctionsImpl.cu:639-669): ov0, ov1 receive overlap0 and overlap1 as
 * complex float [nz][ny][nx] (2*nx*ny*nz floats each, unnormalised inverse
 * transform).  Call after simr_find_modulation and before
 * simr_filter_assemble: the volumes are built in the assembly staging memory. */
int simr_debug_overlap_volumes(simr_ctx* ctx, int dir, int order1, int order2, float k0x, float k0y, float* ov0,
                               float* ov1);

/*
 * Batched complex FFT utility (the engine's own LDS Stockham kernels; replaces
 * the cufftPlan3d/cufftExecC2C pairs at src/cudaSirecon/gpuFunctionsImpl.cu:639-669).
 * data: interleaved complex float [nz][ny][nx], unnormalised, forward = exp(-i),
 * every length 1 to 4096.  Lengths of the form 2^a 3^b 5^c (and 7) run on the
 * transform kernels (simr_fft_size_supported); any other length runs as a
 * direct DFT along that axis through a scratch volume of nx*ny*nz complex
 * values that the library allocates for the call.
 * _dev runs in place on device memory on `stream` (NULL = default stream);
 * with a direct-DFT length it also waits for the stream before returning.
 * _host copies to the current device, transforms and copies back.
 */
int simr_fft3d_dev(float* data, int nx, int ny, int nz, int inverse, void* stream);
int simr_fft3d_host(float* data, int nx, int ny, int nz, int inverse);
int simr_fft_size_supported(int n);
/*
 * simr_fft3d_dev without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Same arguments,
 * lengths and results as simr_fft3d_dev; asynchronous on `stream` in every
 * case.  Its twiddle tables and direct-DFT scratch live in one 64 MiB device
 * array per GPU, held until the process ends: with a length that is not
 * 2^a 3^b 5^c (or 7) the array holds at most simr_fft3d_noalloc_max_elems()
 * complex values (larger arrays are refused with SIMR_ERR_ARG).  Calls that
 * overlap on one GPU share that array and race; serialise them.
 */
int simr_fft3d_dev_noalloc(float* data, int nx, int ny, int nz, int inverse, void* stream);
size_t simr_fft3d_noalloc_max_elems(void);

/*
 * makeotf: PSF -> rotationally averaged OTF (replaces the reference's
 * makeotf program, src/otf/radialft.cpp:82-434, whose FFT is FFTW on the
 * CPU).  Field meanings and defaults follow its command line
 * (radialft.cpp:86-114, 1381-1553).
 */
typedef struct simr_otf_params {
  int32_t nphases;         /* 5; nphases == 1 gives a widefield OTF */
  float k0angle;           /* 1.57 rad (-angle) */
  float linespacing;       /* 0.2 um (-ls) */
  float bead_diameter;     /* 0.12 um (-beaddiam) */
  float na, nimm;          /* 1.4, 1.515 */
  int32_t wavelength;      /* emission, nm (525; MRC input: header iwav1) */
  float dr, dz;            /* pixel sizes, um (0.106, 0.1; MRC input: header ylen, zlen) */
  float background;        /* < 0: estimate each section from its 20-pixel border (TIFF input) */
  int32_t napodize;        /* 10 edge-blend rows; -1: cosine apodization */
  int32_t do_compen;       /* 1: bead-size compensation (-nocompen sets 0) */
  int32_t five_bands;      /* 0: combine each order's re/im bands into one (-5bands sets 1) */
  int32_t conjugate;       /* -conj */
  int32_t dorescale;       /* -rescale: normalise every order by its own origin value */
  int32_t twolens;         /* -2lenses */
  int32_t coherent_bsim;   /* -bessel: skip the out-of-band cleanup */
  int32_t forced_pi_shift; /* -PIshift */
  int32_t ifixz, ifixr;    /* 1, 0 (-ifixkz, -ifixkr) */
  int32_t interpkr[2];     /* -fixorigin kx1 kx2 (0 0: off) */
  int32_t leavekz[3];      /* -leavekz a b c (0 0 0: no cleanup) */
  int32_t reserved[8];
} simr_otf_params;

typedef struct simr_otf_info {
  int32_t nx, ny, nz;      /* PSF geometry (nz = sections / nphases) */
  int32_t norders;         /* (nphases + 1) / 2 */
  int32_t nbands_out;      /* norders, or 2*norders-1 with five_bands */
  int32_t nkr;             /* min(nx, ny)/2 + 1 radial samples */
  float dkr, dkz;          /* OTF pixel sizes, 1/um (1/(nx*dr), 1/(nz*dz)) */
  float xc, yc, zc;        /* fitted bead centre, pixels */
  float phi[SIMR_MAX_ORDERS]; /* side-band phase removed by the re/im combination */
  float scalefactor;       /* 1 / Re(order-0 OTF at the origin) */
  int32_t reserved[8];
} simr_otf_info;

void simr_otf_params_default(simr_otf_params* p);
/* Output size in complex values: nbands_out * nkr * nz. */
size_t simr_otf_elems(const simr_otf_params* p, int nx, int ny, int nsections);
/*
 * psf: nsections = nz * nphases float sections of nx*ny, in Z->Phase order.
 * section_bg: per-section backgrounds (NULL: p->background, or the border
 * estimate when it is negative).  otf: complex float [nbands_out][nkr][nz]
 * (the [kr][kz] layout cudasirecon reads).  band0: optional, receives the
 * separated order-0 band in real space, float [nz][ny][nx] (-gen_order0).
 */
int simr_makeotf(const simr_otf_params* p, const float* psf, int nx, int ny, int nsections, const float* section_bg,
                 float* otf, float* band0, simr_otf_info* info);
/*
 * The I2M OTF of makeotf -I2M (replaces src/otf/radialft.cpp:267-286 reading,
 * 289-296 and determine_center's I2M search 624-664, 306-307, 345-357,
 * 383-386 and 425-430): i2m holds the nz I2M sections [nz][ny][nx].  The
 * last pixel of a section is replaced by its left neighbour when it is
 * exactly -1; the sections are not apodized; the bead centre is the I2M
 * volume's own peak with parabola refinement; one background, the 20-pixel
 * border estimate of section nz/2, is subtracted from every section; no
 * bead-size compensation, modify or cleanup; then p->interpkr's fixorigin
 * and the order-0 rescale.  Only p->dr, dz and interpkr are read.
 * otf: complex float [nkr][nz], nkr = min(nx, ny)/2 + 1; info as for
 * simr_makeotf (nphases 1).
 */
int simr_makeotf_i2m(const simr_otf_params* p, const float* i2m, int nx, int ny, int nz, float* otf,
                     simr_otf_info* info);

/* Synchronise the context stream.  Returns the HIP stream as void*. */
int simr_sync(simr_ctx* ctx);
void* simr_stream(simr_ctx* ctx);

/* Per-stage device timings of the last processed volume, from HIP events on
 * the context stream.  names/ms arrays of capacity cap; returns count. */
int simr_stage_timings(simr_ctx* ctx, const char** names, float* ms, int cap);
int simr_enable_timing(simr_ctx* ctx, int on);
/* Restricts the timing to the stages named `stage` (NULL or "": every stage),
 * so that a timed run carries the events of one kernel only.  Stage names:
 * resample, section_sums, rows_fwd, cols_fwd_y, cols_fwd_z, ov_x, ov_y,
 * ov_zsum, modamp, findk0_peak, filter0 (order 0), filter (side bands),
 * asm_z, asm_y, asm_x. */
int simr_time_stage(simr_ctx* ctx, const char* stage);

/*
 * Serving: a request loop over many stacks of mixed sizes on one GPU (no
 * counterpart in the reference, which reconstructs one stack per process).
 *
 * Optical setups (parameters + OTF) are registered once.  Each request names
 * a setup and a stack shape; up to max_inflight requests run at once, each
 * on its own context (own stream and arena), so the work of several stacks
 * overlaps on the device.  Contexts are cached per (setup, shape) and reused,
 * which keeps the OTFs resident; a request is admitted when a cached idle
 * context of its key exists or a new one fits in hbm_budget bytes (idle
 * contexts of other keys are evicted, least recently used first; a second
 * context of a key whose context is running is made only when it fits
 * without evicting, otherwise the request waits for the running one).
 * Requests are taken in arrival order, skipping over ones that cannot be
 * admitted yet.
 * Time series: a request with series >= 0 and t > 0 waits for t = 0 of its
 * series, then runs on any context with that series' time-0 state, so its
 * result is the one a single sequential context gives.
 */
typedef struct simr_server simr_server;

typedef struct simr_request {
  int32_t setup;           /* id from simr_server_add_setup */
  int32_t nx, ny, nz;      /* stack shape; nImages = nz * ndirs * nphases of the setup */
  const void* raw;         /* the stack in the setup's section order */
  int32_t raw_u16;         /* 1: uint16 pixels, 0: float32 */
  int32_t raw_on_device;   /* 1: raw is a device pointer */
  float* out;              /* receives the zoomed volume [z][y][x] (simr_result_elems floats) */
  int32_t out_on_device;   /* 1: out is a device pointer */
  int32_t series;          /* -1, or the id of the time series this stack belongs to */
  int32_t t;               /* time index within the series */
  int32_t reserved[4];
} simr_request;

typedef struct simr_request_info {
  int32_t status;          /* SIMR_OK or the error code of the failed stage */
  int32_t context_reused;  /* 1: ran on a cached context */
  double t_submit_us, t_start_us, t_end_us; /* steady-clock times of submission, admission, completion */
  size_t out_elems;
  simr_fit fit[SIMR_MAX_DIRS];
  char error[256];
} simr_request_info;

typedef struct simr_server_stats {
  int64_t submitted, completed, failed;
  int64_t contexts_created, contexts_evicted;
  size_t bytes_in_use, bytes_peak;
} simr_server_stats;

int simr_server_create(int device, size_t hbm_budget, int max_inflight, simr_server** out);
/* p: the setup's parameters (nx, ny, nz are taken from each request); otf as for simr_create. */
int simr_server_add_setup(simr_server* srv, const simr_params* p, const float* otf, int* setup_id);
/* Queue a request; returns at once with its ticket.  raw and out must stay valid until it completes. */
int simr_server_submit(simr_server* srv, const simr_request* rq, int64_t* ticket);
/* Block until the request completes; info may be NULL.  Returns the request's status. */
int simr_server_wait(simr_server* srv, int64_t ticket, simr_request_info* info);
int simr_server_stats_get(simr_server* srv, simr_server_stats* st);
/* Waits for every queued request, then frees every context. */
int simr_server_destroy(simr_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* SIMR_H */
