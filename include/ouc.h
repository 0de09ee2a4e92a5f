/*
 * ouc.h -- C ABI of the MI355X elastic P/S-decoupled RTM engine.
 *
 * The reference (gm340/ouc) has no API: its whole path is inlined in main()
 * (P-S-seperating.cu:1658-2493), configured by compile-time #defines
 * (P-S-seperating.cu:14-26) and locals (P-S-seperating.cu:1685-1688,
 * 2113-2114), and writes raw float32 image files (P-S-seperating.cu:2379-2467).
 * Each entry point below replaces one seam of that main(); the comment on each
 * names the reference lines it stands in for.
 *
 * Conventions
 *  - Plain C types only; no exceptions cross the ABI.  Every call returns an
 *    ouc_status; ouc_last_error() gives the message of the last failure.
 *  - Host arrays are caller owned.  Interior arrays are nx*nz floats, x-major
 *    (index ix*nz + iz, z fastest), the order the reference writes its images in
 *    (P-S-seperating.cu:2383-2387).
 *  - Coordinates are interior-relative: padded = interior + npml, the
 *    reference's convention (Sx = L + 5l, Sz = Z_receive = L at
 *    P-S-seperating.cu:2113-2114).
 *  - The library owns all device memory.  Calls block until results reach the
 *    host.  One host thread per context; contexts are independent.
 */
#ifndef OUC_H
#define OUC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ouc_status {
    OUC_OK = 0,
    OUC_INVALID_ARG = 1,
    OUC_OOM = 2,
    OUC_DEVICE = 3,
    OUC_NUMERIC_FALLBACK = 4, /* learned boundary diverged; CPML result returned */
    OUC_IO = 5,
    OUC_UNSUPPORTED = 6
} ouc_status;

/* Time-index semantics of the source-wavefield reconstruction.
 * FAITHFUL reproduces the reference: the velocity boundary strips restored at
 * backward step t hold V^{t+1} (saved after forward_u, P-S-seperating.cu:2264
 * then 2272, restored at 2337), and the imaging condition correlates the source
 * P velocity two steps ahead (d_Vp*_s after reshot_u, P-S-seperating.cu:964-967
 * and 2361).  EXACT restores V^t and correlates V^{t+1}, which makes the
 * reconstruction the exact time reverse of the forward run. */
enum { OUC_SEM_FAITHFUL = 0, OUC_SEM_EXACT = 1 };

/* Storage of the boundary strips (P-S-seperating.cu:1852-1893 allocate them in
 * FP32).  FP16 stores, per ring cell, the residual of each slot against the
 * previous slot's decoded value, with the previous quantization error fed
 * back, as binary16 with a power-of-two scale per 64-value chunk; every 32nd
 * slot keeps an FP32 keyframe, and the reconstruction decodes backwards one
 * slot per step.  The error the reconstruction sees is then a first time
 * difference (small at the wavefield's frequencies) of a quantity about the
 * size of one step's change times 2^-12, rather than white noise of 2^-12 of
 * the value.  Compute stays FP32. */
enum { OUC_STRIP_FP32 = 0, OUC_STRIP_FP16 = 1 };

/* Absorbing boundary.  CPML is the reference's (P-S-seperating.cu:2067-2105).
 * LEARNED replaces most of the CPML pad by a thin band driven by a small
 * convolutional network; LEARNED_FALLBACK does the same and falls back to CPML
 * when the band's energy grows. */
enum { OUC_BOUNDARY_CPML = 0, OUC_BOUNDARY_LEARNED = 1, OUC_BOUNDARY_LEARNED_FALLBACK = 2 };

/* Image filter.  STENCIL is the reference's 5-point Laplace kernel
 * (P-S-seperating.cu:1457-1469); FFT5 applies the same operator as a rocFFT
 * R2C/C2R filter with the 5-point discrete symbol; FFT_CONT uses -|k|^2. */
enum { OUC_LAP_STENCIL = 0, OUC_LAP_FFT5 = 1, OUC_LAP_FFT_CONT = 2 };

/* Kernel structure.  PLAIN keeps the reference's kernel split (separate stress
 * pass, velocity pass, strip copy, reconstruction pair and imaging kernels,
 * P-S-seperating.cu:2258-2364).  FUSED runs one single-pass kernel per
 * wavefield per step with strip I/O and the imaging condition folded in.
 * SPLIT_PML is the reference's other formulation, defined but never launched
 * by its main(): first-order velocity-stress with split-field PML damping
 * (stress P-S-seperating.cu:487-555, velocity 390-485; 12th-order
 * coefficients to five digits at N = 6), with the context's CPML profiles.
 * ouc_forward_shot runs it (no mute).  ouc_rtm_shot and ouc_rtm_batch run its
 * backward half too (receiver run with velocity_backward_propagation /
 * stress_backward_propagation, P-S-seperating.cu:728-887; source
 * reconstruction with velocity_ / stress_wavefield_reconstruction,
 * P-S-seperating.cu:1066-1229, from FP32 strips), with EXACT time alignment
 * only; the FWI calls return OUC_UNSUPPORTED for it. */
enum { OUC_ENGINE_PLAIN = 0, OUC_ENGINE_FUSED = 1, OUC_ENGINE_SPLIT_PML = 2 };

/* Run parameters.  Replaces the #defines L, N, NT, Xn, Zn, R
 * (P-S-seperating.cu:15-23) and the locals dt, dx, dz
 * (P-S-seperating.cu:1685-1688).  Xn = nx + 2*npml, Zn = nz + 2*npml. */
typedef struct ouc_params {
    int32_t nx, nz;          /* interior cells */
    int32_t npml;            /* L: absorbing pad cells on each side */
    int32_t half_order;      /* N: stencil half-order, 1..8 (reference: 6) */
    int32_t nt;              /* time steps (reference NT = 6000) */
    float dx, dz, dt;        /* metres, seconds; the mute and the Laplacian use dx as dh */
    double R;                /* CPML reflection coefficient (reference 1e-3) */
    int32_t semantics;       /* OUC_SEM_* */
    int32_t strip_precision; /* OUC_STRIP_* */
    int32_t boundary;        /* OUC_BOUNDARY_* */
    int32_t laplacian;       /* OUC_LAP_* */
    int32_t engine;          /* OUC_ENGINE_* */
    int32_t band;            /* learned-boundary band width in cells (0: default 8) */
} ouc_params;

/* Elastic model on the interior, x-major nx*nz.  The pad is filled by edge
 * replication inside the library, as xiangbian() does
 * (P-S-seperating.cu:37-119).  rho is accepted and unused: the live physics has
 * no density (P-S-seperating.cu:373-378). */
typedef struct ouc_model {
    const float *vp;
    const float *vs;
    const float *rho; /* may be NULL */
} ouc_model;

/* One shot.  Replaces Sx, Sz, Z_receive (P-S-seperating.cu:2113-2114) and the
 * Ricker wavelet (P-S-seperating.cu:1954-1965).  Receivers are every interior
 * column of row rec_z, as the reference records every column of its receiver
 * row (P-S-seperating.cu:307-311) and injects the interior ones
 * (P-S-seperating.cu:1477-1481). */
typedef struct ouc_shot {
    int32_t sx, sz;          /* source cell */
    int32_t rec_z;           /* receiver row */
    float fm;                /* Ricker peak frequency, Hz; also sets the mute's t0 */
    const float *wavelet;    /* optional nt samples; NULL: Ricker(fm) */
    int32_t mute;            /* 1: direct-wave mute of the records (P-S-seperating.cu:1634-1657) */
    int32_t nt;              /* steps of this shot, 0 = the context's nt.  Fused RTM only: in a batch
                                a shorter shot starts late (zero wavelet before its start) and images
                                only its own steps, so its image equals a run of nt steps alone */
} ouc_shot;

/* Per-call timing, measured with HIP events on the stream the kernels run on. */
typedef struct ouc_timings {
    double total_ms;         /* whole call, device side */
    double forward_ms;       /* forward (source) time loop */
    double backward_ms;      /* receiver + reconstruction + imaging loop */
    double image_ms;         /* normalisation + Laplacian + stack */
    int64_t step_launches;   /* stencil-step kernel launches in the two loops */
    double step_kernel_ms;   /* summed duration of the dominant step kernel */
    int64_t step_kernel_launches;
    double cell_updates;     /* padded-cell x step x wavefield updates performed */
    double canonical_bytes;  /* reference two-pass compulsory bytes for the same work */
    double own_bytes;        /* compulsory bytes of the kernels this engine ran */
    int32_t fallback_shots;  /* shots re-run with CPML after the learned boundary diverged */
    int32_t reserved;
    double boundary_ratio;   /* learned boundary: largest |V| in the pad over the largest |V| in the
                                band of the same width inside it, over shots and wavefields */
} ouc_timings;

typedef struct ouc_ctx ouc_ctx;

/* Build the padded model, CPML profiles and device state for one grid.
 * Replaces the setup in main(): allocations (P-S-seperating.cu:1782-1952),
 * model and padding (P-S-seperating.cu:1976-2065), CPML damping
 * (P-S-seperating.cu:2067-2105).  device: HIP ordinal. */
ouc_status ouc_create(const ouc_params *p, const ouc_model *m, int32_t device, ouc_ctx **out);

/* Forward-model one shot and return its seismograms.  Replaces the forward
 * time loop (P-S-seperating.cu:2258-2292: forward_s, forward_u with the
 * recording at 307-311).  seis: caller-owned [6][nx][nt] floats, components
 * vx, vz, vpx, vpz, vsx, vsz recorded after each step's velocity update at row
 * rec_z; receiver-major like record_vx[i*NT+t].  The P and S parts come from the
 * div/curl split (P-S-seperating.cu:291-294).  Unmuted unless shot->mute. */
ouc_status ouc_forward_shot(ouc_ctx *c, const ouc_shot *s, float *seis);

/* One RTM shot: forward run with boundary-strip saving, direct-wave mute,
 * backward loop (receiver propagation with record injection, source
 * reconstruction from the strips, PP/PS imaging), normalisation and Laplacian.
 * Replaces one pass of the shot loop (P-S-seperating.cu:2108-2376).
 * rec_in: optional observed [2][nx][nt] (vx, vz), receiver-major; NULL uses the
 * shot's own modelled records, as the reference does (P-S-seperating.cu:2293-2310).
 * Images are interior nx*nz x-major; any output pointer may be NULL. */
ouc_status ouc_rtm_shot(ouc_ctx *c, const ouc_shot *s, const float *rec_in,
                        float *img_pp, float *img_ps, float *img_pp_lap, float *img_ps_lap);

/* Batched RTM over many shots on one grid: the shots run together, one launch
 * per time step covering every shot, with the time loops captured in HIP
 * graphs.  Returns the stack (sum over shots, in shot order) of the Laplacian
 * images, which replaces the stacking pass (P-S-seperating.cu:2403-2443).
 * per_shot_ms: optional nshots floats, each shot's completion time from the
 * start of the call. */
ouc_status ouc_rtm_batch(ouc_ctx *c, const ouc_shot *shots, int32_t nshots,
                         float *stack_pp_lap, float *stack_ps_lap, float *per_shot_ms);

/* One job's shots on several contexts of the same grid at once (fused
 * engine): the shots are cut into batches of one size, batch j runs on
 * context j mod nctx, and every context works on its own stream from its own
 * host thread, so the band kernels of one batch run beside the core kernels of
 * another.  The batch size is what nctx equal shares of the free device
 * memory hold, within each context's ouc_set_batch_capacity (pinned for the
 * call).  The contexts must share the device and the parameters (not the
 * tile); they are meant to hold the same model.  Returns the stack of the
 * Laplacian images summed in shot order on the host: bit for bit what
 * ouc_rtm_batch gives for the same shots on one context (shots that fall back
 * to CPML are added in shot order here, after the others there).
 * per_shot_ms: optional, each shot's batch completion time from the start of
 * the call.  Timings of the call go to ctxs[0].  At most 8 contexts.  No
 * counterpart in the reference, which runs its shots one after another (the
 * shot loop, P-S-seperating.cu:2108-2401, and its stack, 2403-2443). */
ouc_status ouc_rtm_concurrent(ouc_ctx *const *ctxs, int32_t nctx, const ouc_shot *shots, int32_t nshots,
                              float *stack_pp_lap, float *stack_ps_lap, float *per_shot_ms);

/* Batched RTM over shots on several grids at once (fused engine): context i
 * (its own grid and model) runs shots[i][0 .. nshots[i]).  Every kernel of a
 * time step is one launch over the shots of all contexts, whatever their grid
 * sizes, with the time loops captured in HIP graphs; contexts with fewer steps
 * join the loops late and end with the others.  Each context's images equal
 * what ouc_rtm_batch gives for its own shots.  stack_pp_lap[i], stack_ps_lap[i]
 * receive context i's stacks (either array, or an entry, may be NULL).  The
 * contexts must share the device, half-order (5 or 6), boundary kind, strip
 * precision and semantics; at most 8 contexts.  With LEARNED_FALLBACK,
 * diverged shots run again on CPML and the call returns
 * OUC_NUMERIC_FALLBACK with valid stacks.  No counterpart in the reference,
 * whose shots all share one grid (the shot loop, P-S-seperating.cu:2108-2401,
 * and its stack, 2403-2443). */
ouc_status ouc_rtm_grouped(ouc_ctx *const *ctxs, int32_t nctx, const ouc_shot *const *shots,
                           const int32_t *nshots, float *const *stack_pp_lap, float *const *stack_ps_lap);

/* Profiling mode: when on, the forward loop brackets a sample of the dominant
 * step kernel's launches with HIP events and reports their summed duration in
 * ouc_timings.step_kernel_ms.  No counterpart in the reference, which never
 * times itself (P-S-seperating.cu:1663). */
ouc_status ouc_set_profiling(ouc_ctx *c, int32_t on);

/* HIP graphs (default on): the fused engine captures each time loop of a
 * batch (forward, then receiver + reconstruction) with stream capture the
 * first time a batch size is run and replays the graph on later calls.  In
 * the captured backward loop the reconstruction of each step runs on a second
 * stream beside the receiver step (they share no data within a step) and the
 * two are joined every step.  While profiling is on, the forward loop (which
 * carries the sampled kernel timing) runs eagerly.  Off: every loop runs
 * eagerly on one stream.  No counterpart in the reference, which launches
 * every kernel from the host loop (P-S-seperating.cu:2258-2364). */
ouc_status ouc_set_graphs(ouc_ctx *c, int32_t on);

/* Shots per fused launch: ouc_rtm_batch runs at most `shots` shots together
 * (0: as many as 85% of free device memory holds), so contexts that share a
 * GPU can split its memory.  bytes_per_shot (optional) returns what one shot
 * of a batch holds on the device. */
ouc_status ouc_set_batch_capacity(ouc_ctx *c, int32_t shots, double *bytes_per_shot);

/* Tile of the fused engine's interior kernel: x columns each 256-lane block
 * streams and the columns of loads kept in flight (1..3).  Results do not
 * depend on the tile (each cell's arithmetic is the same); speed does.  No
 * counterpart in the reference, whose launch shape is fixed
 * (P-S-seperating.cu:1978-1979). */
typedef struct ouc_tile {
    int32_t core_cols;
    int32_t prefetch;
} ouc_tile;
ouc_status ouc_set_tile(ouc_ctx *c, const ouc_tile *t);
ouc_status ouc_get_tile(ouc_ctx *c, ouc_tile *t);

/* Autotuner: times the interior kernel on this context's grid with nshots
 * shots, `reps` launches per candidate (core_cols {8, 16, 32, 64, 128, W}
 * x prefetch {1, 2, 3}: 18 candidates, ms_per_launch[18] optional), selects
 * the fastest and returns it in *best.  W is the width at which one round of
 * the blocks the device holds at once covers the launch. */
ouc_status ouc_autotune(ouc_ctx *c, int32_t nshots, int32_t reps, ouc_tile *best, float *ms_per_launch);

/* Streaming-rate probe (no counterpart in the reference; a measurement aid):
 * times `reps` launches of a kernel in which every thread reads nin arrays
 * and writes nout arrays of n floats at its own index (1..16 each), the
 * access pattern of the core kernel's state streams, and returns the mean
 * launch time.  The last nshared (< nin) inputs hold only m floats and are
 * read at index mod m, as the core kernel reads the model plane that all shots
 * share.  width (1, 2 or 4) is the number of consecutive floats each thread
 * moves per array (1: the core kernel's one cell per lane); width -2 keeps one
 * value per thread for the arithmetic and moves 256-float chunks as float2
 * loads and stores transposed through LDS, width -3 moves two arrays per
 * float2 instruction within a wave and redistributes them with cross-lane
 * permutes (no shared inputs for either).  n is a multiple of 256 and m of 4.
 * Device memory is held for the call only. */
ouc_status ouc_stream_probe(int32_t device, int32_t nin, int32_t nout, int64_t n, int32_t nshared, int64_t m,
                            int32_t width, int32_t reps, float *ms_per_launch);

/* gfx architecture name of a HIP device (e.g. "gfx950"), for tuning caches. */
ouc_status ouc_device_arch(int32_t device, char *buf, int32_t len);

/* Elastic FWI gradient of a set of shots, summed over the shots (fused engine,
 * CPML boundary, every shot at the context's nt).  Replaces the gradient pass
 * the reference declares but never launches: the adjoint residual
 * (adsource, P-S-seperating.cu:1484-1498: modelled minus observed on every
 * receiver and step, no mute) drives the receiver wavefield, and gradfun
 * (P-S-seperating.cu:1518-1575) accumulates at every backward step t
 *   grad_vp -= 2 vp (div u) theta_r
 *   grad_vs -= 2 vs (omega_r (duxdz - duzdx) + duzdx_r duzdx + duzdz_r duzdz
 *                    + duxdz_r duxdz + duxdx_r duxdx)
 *   illum   += ux^2 + uz^2
 * with u the reconstructed source velocity at level t+1 (the level the EXACT
 * imaging condition uses) and the receiver stress fields of step t.  The
 * derivatives of u use the engine's stencil: the reference's hard-coded
 * 12th-order duxdz pairs f[j+4] with f[j-6] (P-S-seperating.cu:1556), which is
 * not carried over.  The live physics has no density, so the reference's factor
 * 2 P vp is 2 vp.  observed: [nshots][2][nx][nt] (vx, vz), receiver-major;
 * grad_vp, grad_vs, illum: interior nx*nz x-major, any may be NULL; misfit
 * (may be NULL): 0.5 * sum of squared residuals over all shots. */
ouc_status ouc_fwi_gradient(ouc_ctx *c, const ouc_shot *shots, int32_t nshots, const float *observed,
                            float *grad_vp, float *grad_vs, float *illum, double *misfit);

/* Model update of one array: out = v - step * 0.95^iter * grad, evaluated in
 * double and rounded to float, as revise_model_end (P-S-seperating.cu:1500-1517,
 * step a_2 = 30 at P-S-seperating.cu:25) does for Vp and Vs on the interior.
 * Host arrays of n floats; device memory is held for the call only. */
ouc_status ouc_model_update(int32_t device, int64_t n, const float *v, const float *grad, double step, int32_t iter,
                            float *out);

/* Model dump: dir/Vp.dat and dir/Vs.dat, the interior as raw little-endian
 * float32, x-major (the layout of the reference's model dump, kept commented
 * out at P-S-seperating.cu:2042-2063).  Creates dir; either array may be NULL. */
ouc_status ouc_write_model_files(const char *dir, int32_t nx, int32_t nz, const float *vp, const float *vs);

/* One shot gather of one component as a SEG-Y revision 1 file.  The
 * reference records seismograms (P-S-seperating.cu:307-311) but never writes
 * them.  Layout: a 3200-byte EBCDIC textual header, a 400-byte binary header
 * (sample format 5 = IEEE float, fixed-length traces), then per receiver a
 * 240-byte trace header (sequence numbers, field record = shot, offset,
 * source/receiver X, depths as elevations, coordinate scalar -100) and nt
 * big-endian float samples.  traces: [ntraces][nt] floats, receiver-major
 * like a component of the seismogram arrays. */
typedef struct ouc_segy_gather {
    int32_t shot;       /* field record number */
    int32_t ntraces;    /* receivers, one trace each */
    int32_t nt;         /* samples per trace, at most 32767 */
    float dt;           /* sample interval, seconds (1 us .. 32.767 ms) */
    float sx_m, sz_m;   /* source position, metres */
    float rec_x0_m, rec_dx_m, rec_z_m; /* receiver k at x = rec_x0 + k rec_dx, depth rec_z */
    char component[8];  /* written to the textual header, e.g. "VX" */
} ouc_segy_gather;
ouc_status ouc_write_segy(const char *path, const ouc_segy_gather *g, const float *traces);

/* Reads a file written by ouc_write_segy (or any fixed-length, format-5,
 * revision-1 file without extended textual headers).  With traces NULL,
 * returns the sizes only; otherwise fills [ntraces][nt] floats, capacity
 * being the buffer's length in floats.  dt may be NULL. */
ouc_status ouc_read_segy(const char *path, int32_t *ntraces, int32_t *nt, float *dt, float *traces, int64_t capacity);

/* Timing of the last forward/RTM/batch call on this context. */
ouc_status ouc_get_timings(const ouc_ctx *c, ouc_timings *t);

/* Write one shot's Laplacian images in the reference's file format and names:
 * <dir>/image_PP_lap_<nx>_<nz>_<shot>.dat and image_PS_lap_..., raw
 * little-endian float32, x-major interior (P-S-seperating.cu:2379-2400).
 * shot < 0 writes the stack names All_image_PP_lap.dat / All_image_PS_lap.dat
 * (P-S-seperating.cu:2446-2467).  Creates dir if missing. */
ouc_status ouc_write_image_files(const char *dir, int32_t nx, int32_t nz, int32_t shot,
                                 const float *pp_lap, const float *ps_lap);

/* Scaled-FP16 strip codec on the GPU, for bit-exact testing against a host
 * implementation: encodes `steps` consecutive slots of one field, x as
 * [steps][n] floats (host), in 64-value chunks exactly as the strip store
 * does (predictive coding with error feedback, encoder state zero at the
 * first slot, the last slot a keyframe).  Returns the binary16 residual bits
 * [steps][n], each chunk's power-of-two exponent [steps][ceil(n/64)] and, if
 * `decoded` is not NULL, the values [steps][n] decoded backwards as the
 * reconstruction decodes them.  No counterpart in the reference, whose strips
 * are FP32 (P-S-seperating.cu:1852-1893). */
ouc_status ouc_strip_codec_fp16(int32_t device, const float *x, int64_t n, int32_t steps, uint16_t *half_bits,
                                int8_t *exponents, float *decoded);

/* Learned absorbing boundary (boundary LEARNED / LEARNED_FALLBACK; no
 * counterpart in the reference, whose only boundary is the CPML of
 * P-S-seperating.cu:2067-2105).  The outer `band` cells run second-order
 * staggered differences (with a weak sponge, see ouc_set_learned_damping); a three-layer 3x3 convolutional network (5 -> 16 ->
 * 16 -> 2 channels, linear, BF16 weights, FP32 accumulation) sets the outermost
 * ring from the local velocity increment and stress.  weights: FP32 masters,
 * n = 3312 (W1 [16][5][3][3], W2 [16][16][3][3], W3 [2][16][3][3]), rounded to
 * BF16 here; NULL holds the ring at zero (the pad without the network).
 * A context with the learned boundary starts without a network. */
ouc_status ouc_set_learned_boundary(ouc_ctx *c, const float *weights, int32_t n);

/* Weak sponge in the learned pad: the velocity increment of pad cells at depth
 * d is scaled by 1 - beta ((band - d) / band)^2 each step (0 <= beta < 1;
 * default 0.01).  It removes the slow growth the one-step-trained ring shows
 * after several thousand steps, at about 1-2 dB of reflection. */
ouc_status ouc_set_learned_damping(ouc_ctx *c, float beta);

/* The composed 7x7 filter [2][5][7][7] the ring kernel applies (test hook). */
ouc_status ouc_get_learned_filter(ouc_ctx *c, float *taps);

/* Forward run of one shot that copies V (vx then vz) over the padded-grid
 * rectangle [x0, x0+w) x [z0, z0+h) after each listed step (increasing) into
 * out[n][2][w][h].  This is how the engine generates the wavefields its
 * learned boundary trains on; no counterpart in the reference. */
ouc_status ouc_forward_snapshots(ouc_ctx *c, const ouc_shot *s, int32_t n, const int32_t *steps, int32_t x0,
                                 int32_t z0, int32_t w, int32_t h, float *out);

/* RTM of one shot that copies the reconstructed source V^t (vx then vz) over
 * the padded-grid rectangle at each listed time (decreasing, in [0, nt-2])
 * into out[n][2][w][h].  Against ouc_forward_snapshots (V^t is the snapshot
 * after forward step t-1) it measures the time reversibility of the
 * reconstruction (reshot_u / reshot_s, P-S-seperating.cu:890-1042). */
ouc_status ouc_recon_snapshots(ouc_ctx *c, const ouc_shot *s, int32_t n, const int32_t *steps, int32_t x0,
                               int32_t z0, int32_t w, int32_t h, float *out);

/* Training of the learned-boundary network on the GPU (no context).  Windows
 * x [M][5][H][W], targets y [M][2][H][W], mask [M][H][W] (1 where the loss
 * counts), weights as in ouc_set_learned_boundary.  Loss = sum mask (out -
 * y)^2 / (2 sum mask); the forward pass uses the BF16-rounded weights, and the
 * gradient is taken with respect to them. */
ouc_status ouc_lb_forward(int32_t device, int32_t M, int32_t H, int32_t W, const float *x, const float *weights,
                          float *out);
/* ouc_lb_forward without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  x, weights and out are DEVICE pointers with the
 * layouts of ouc_lb_forward; the BF16 weight copy and the hidden activations
 * live in a static per-GPU scratch array (16 MiB + 8 KiB, held until the
 * process ends), so M*H*W is at most 131072 (else OUC_INVALID_ARG).  The
 * launches are queued on `stream` (a hipStream_t; NULL for the null stream)
 * and the call returns without waiting.  Calls that overlap on one GPU share
 * the scratch and race: order them on one stream. */
ouc_status ouc_lb_forward_device(int32_t device, int32_t M, int32_t H, int32_t W, const float *x,
                                 const float *weights, float *out, void *stream);
ouc_status ouc_lb_loss_grad(int32_t device, int32_t M, int32_t H, int32_t W, const float *x, const float *y,
                            const float *mask, const float *weights, float *loss, float *grad);
/* Adam (beta 0.9, 0.999, eps 1e-8) over `epochs` full-batch steps; weights
 * in/out; history: optional per-epoch losses. */
ouc_status ouc_lb_train(int32_t device, int32_t M, int32_t H, int32_t W, const float *x, const float *y,
                        const float *mask, int32_t epochs, float lr, float *weights, float *history);

/* Message of the last failed call on c (or of the last failed ouc_create when
 * c is NULL).  Never NULL. */
const char *ouc_last_error(const ouc_ctx *c);

void ouc_destroy(ouc_ctx *c);

#ifdef __cplusplus
}
#endif

#endif /* OUC_H */
