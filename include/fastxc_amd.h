/*
 * fastxc_amd.h -- C ABI of the MI355X (gfx950) FastXC kernels.
 *
 * The reference FastXC has no library boundary: Python runs the four stage
 * executables (fastxc/runtime/command_runner.py:147-155) and everything
 * crosses as files.  This header is the thin C ABI the replacement stage
 * binaries call, one entry point per reference hot-path routine it replaces
 * (cited per declaration).  Conventions:
 *   - every function returns 0 on success, non-zero on error; the message is
 *     available from fxc_last_error() (thread-local);
 *   - buffers are caller-owned device pointers (complex data is interleaved
 *     float re/im); `stream` is a hipStream_t or NULL for the default stream;
 *   - work is enqueued on `stream` and is complete when the stream is.
 */
#ifndef FASTXC_AMD_H
#define FASTXC_AMD_H

/* Status codes returned by every entry point. */
#define FXC_OK 0
#define FXC_ERR_ARG 1         /* invalid argument or shape */
#define FXC_ERR_HIP 2         /* a HIP runtime call failed */
#define FXC_ERR_ALLOC 3       /* device allocation failed */
#define FXC_ERR_UNSUPPORTED 4 /* valid but not supported (length, window count) */

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed call on this thread. */
const char *fxc_last_error(void);
/* Frees cached FFT plans (twiddle tables) of every device. */
void fxc_release_plans(void);

/* ---------------------------------------------------------------- FFT --
 * Batched transforms on the workgroup LDS FFT (lengths whose n (complex) or
 * n/2 (real) factor into 2,3,5,7,23,29 and fit 20480 complex values).
 * Replace the cuFFT plans of native/sac2spec/device_memory.cu:23-32,
 * native/xc/memory/memory.cu:478-483, native/pws/worker/gpu.cu:197-227 and
 * native/tfpws/core/tfpws_workspace.cu:190-235.  Unnormalised; `scale`
 * multiplies every output. */
int fxc_fft_c2c(const float *d_in, long long in_dist, float *d_out, long long out_dist, int n,
                long long batch, int inverse, float scale, void *stream);
/* Real forward transform: in_len samples per row (zero-padded to n); writes
 * bins 0..n/2. */
int fxc_fft_r2c(const float *d_in, long long in_dist, int in_len, float *d_out,
                long long out_dist, int n, long long batch, float scale, void *stream);
/* Real inverse transform of bins 0..n/2; writes the first out_len samples
 * (imaginary parts of bins 0 and n/2 are ignored). */
int fxc_fft_c2r(const float *d_in, long long in_dist, float *d_out, long long out_dist,
                int out_len, int n, long long batch, float scale, void *stream);

/* ------------------------------------------------------------- xc_fast --
 * Windowed cross-correlation of one timestamp.
 * Spectra layout: [nstep][nsta][nspec] complex64, the StepPack v3 payload
 * order (native/xc/io/input.cu:46-67).  Output: [npairs][2*half_cc+1] float,
 * row t for task t = (src[t], rec[t]) in the caller's order. */
#define FXC_PREC_FP32 0   /* fp32 vector FMA (reference arithmetic) */
#define FXC_PREC_BF16X3 1 /* split-bf16 hi+lo MFMA, fp32 accumulation */
#define FXC_PREC_FP16 2   /* fp16 MFMA with per-station power-of-2 scale */
#define FXC_PREC_PLAIN 3  /* the reference's step order: per-window fp32 stack RMW */
#define FXC_PREC_FP16X3 4 /* split-fp16 hi+lo MFMA (per-station power-of-2 scale), fp32 acc. */

typedef struct fxc_xc_plan fxc_xc_plan;

/* nfft = 2(nspec-1); half_cc = llround(max_lag/dt) clamped to nspec-1, as
 * finalize_shape (native/xc/io/input.cu:1033-1054). */
int fxc_xc_plan_create(fxc_xc_plan **plan, int nspec, int nstep, float dt, int half_cc,
                       int precision);
int fxc_xc_plan_destroy(fxc_xc_plan *plan);
int fxc_xc_plan_info(const fxc_xc_plan *plan, int *nfft, int *half_cc, int *cc_size);
/* Bounds the cross-spectrum workspace (rows of nspec complex); 0 = all pairs. */
int fxc_xc_set_chunk_rows(fxc_xc_plan *plan, long long rows);
/* Task list of the job: src/rec station indices (host arrays).  Replaces the
 * index lists of build_tasks_for_job (native/xc/worker/scheduler.cu:65-152). */
int fxc_xc_set_pairs(fxc_xc_plan *plan, int nsta, const int *h_src, const int *h_rec,
                     long long npairs);
int fxc_xc_pairs_info(const fxc_xc_plan *plan, long long *ntiles, long long *nchunks,
                      long long *max_chunk_rows);
/* Windowed stack only: d_xspec[t][k] = (1/nstep) sum_w conj(S[w][src][k]) S[w][rec][k].
 * Replaces the per-window accumulateStepXc2DKernel loop
 * (native/xc/worker/async_workers.cu:66-83, kernels/cuda.kernels.cu:16-36). */
int fxc_xc_cross_spectra(fxc_xc_plan *plan, const float *d_spec, float *d_xspec, void *stream);
/* Lag synthesis only: (-1)^k shift, C2R(nfft), x 1/(nfft*dt), lags
 * [nfft/2-H, nfft/2+H] of each cross-spectrum row.  Replaces
 * finalize_xc_batch (native/xc/worker/executor.cu:6-45). */
int fxc_xc_lags(fxc_xc_plan *plan, const float *d_xspec, long long nrows, float *d_cc,
                void *stream);
/* Per-kernel timing with HIP events recorded on the launch stream (no
 * reference counterpart; used by the benchmark).  get_timing waits for the
 * events and returns milliseconds and launch counts for the split,
 * contraction and lag-synthesis kernels since the last set_timing. */
int fxc_xc_set_timing(fxc_xc_plan *plan, int enable);
int fxc_xc_get_timing(fxc_xc_plan *plan, double *ms3, long long *launches3);
/* Spec mode (no reference counterpart: the reference contracts all nspec
 * bins, native/xc/worker/async_workers.cu:66-83): the windowed stack is
 * formed only for the 16-bin chunks that hold bins k_lo..k_hi, and every
 * other bin of the cross-spectrum is 0 before the lag synthesis.  For spectra
 * whose whitening band lies inside [k_lo, k_hi] the lags differ from the
 * parity result by the energy the spectra hold outside the band.  k_lo = 0,
 * k_hi = nspec-1 restores the parity mode (the default).  Not for PLAIN. */
int fxc_xc_set_band(fxc_xc_plan *plan, int k_lo, int k_hi);
/* Stack + lag synthesis for every task (execute_row_batch_job,
 * native/xc/worker/async_workers.cu:18-86, minus file I/O).  With one
 * window at nfft 37044 or at a four-step lag length such as 74088 (every
 * bin, any form but PLAIN) there is no sum over windows: the stack
 * conj(S[src]) S[rec] is formed in FP32 by the lag synthesis as it reads
 * the two spectra, whatever the plan's operand form. */
int fxc_xc_correlate(fxc_xc_plan *plan, const float *d_spec, float *d_cc, void *stream);

/* Uneven time segments in one set of launches (split-bf16 form only).
 * Segment s has nsta[s] stations and nstep[s] <= 128 windows; its pairs are
 * (i, j), i < j, row-major, and the output rows [ntasks][2*half_cc+1] are the
 * segments' pairs one segment after another.  d_spec: the segments' spectra
 * [nstep][nsta][nspec] complex back to back; d_valid: null or one byte per
 * station-window ([nstep][nsta] per segment, back to back), 0 = a gap, which
 * enters the stack as zeros.  Gap mode KEEP divides by the segment's window
 * count, as the reference does for zero-filled windows
 * (native/xc/worker/async_workers.cu:66); EXCLUDE divides each pair by the
 * windows in which both stations are valid (0 of them: a zero row).
 * Per segment the result is fxc_xc_correlate's for that segment alone.
 * max_rows bounds the cross-spectrum workspace (0 = no bound); the plans
 * of a batch run one after another on the caller's stream and share one
 * workspace and one set of split operands, sized for the largest plan.
 * No reference counterpart: the reference runs one timestamp's job at a
 * time (native/xc/worker/async_workers.cu:18-86). */
#define FXC_GAPS_KEEP 0
#define FXC_GAPS_EXCLUDE 1
typedef struct fxc_xc_batch fxc_xc_batch;
int fxc_xc_batch_create(fxc_xc_batch **batch, int nspec, float dt, int half_cc, int precision,
                        int nseg, const int *nsta, const int *nstep, int gap_mode,
                        long long max_rows);
int fxc_xc_batch_destroy(fxc_xc_batch *batch);
int fxc_xc_batch_info(const fxc_xc_batch *batch, long long *ntasks, long long *spec_elems,
                      long long *valid_elems, int *ngroups, int *grouping);
/* grouping 0: every segment in one plan (operand depth of the longest
 * segment); 1: one plan per operand depth, the window count rounded up to
 * the next compiled depth (8, 16, 24, 32, 40, 48, 64, 96, 128). */
int fxc_xc_batch_set_grouping(fxc_xc_batch *batch, int grouping);
/* Row cap of the shared cross-spectrum workspace (0 = no cap): the row
 * chunk every persistent contraction and lag launch of the batch covers. */
int fxc_xc_batch_set_rows(fxc_xc_batch *batch, long long max_rows);
int fxc_xc_batch_correlate(fxc_xc_batch *batch, const float *d_spec, const unsigned char *d_valid,
                           float *d_cc, void *stream);
/* Times both groupings on these buffers (reps runs each, after a warm-up)
 * and keeps the faster; ms[2] gets the mean ms per run of each. */
int fxc_xc_batch_tune(fxc_xc_batch *batch, const float *d_spec, const unsigned char *d_valid,
                      float *d_cc, int reps, void *stream, double *ms);
/* Captures one run into a HIP graph bound to these buffers (stream must not
 * be the null stream); replay launches it again on any stream. */
int fxc_xc_batch_capture(fxc_xc_batch *batch, const float *d_spec, const unsigned char *d_valid,
                         float *d_cc, void *stream);
int fxc_xc_batch_replay(fxc_xc_batch *batch, void *stream);

/* ------------------------------------------------------------ sac2spec --
 * Per station-window preprocessing, normalisation, whitening and output
 * spectrum.  Replaces ComputeWorkerBatch (native/sac2spec/worker/compute.cu:8-272)
 * with the plan values main() derives (native/sac2spec/cuda.main.cu:190-296). */
typedef struct fxc_s2s_config {
  int npts;         /* samples per trace (llround(sac_len/delta)) */
  int seg;          /* window length (2,3,5,7-smooth, even) */
  int shift;        /* window shift in samples */
  int nfft_filter;  /* prefilter FFT length (seg + filter padding, even smooth) */
  int nfft_out;     /* output FFT length; nspec = nfft_out/2 + 1 */
  int num_ch;       /* channels per station group (rows of one frame) */
  float delta;      /* sampling interval (s) */
  float freq_low;   /* broad-band low corner (filter 0) */
  int idx[4];       /* whitening taper bins on the seg grid */
  int whiten;       /* 0 none, 1 before, 2 after, 3 both normalisation */
  int normalize;    /* 0 none, 1 RUN-ABS-MF, 2 one-bit, 3 RUN-ABS */
  int phase_only;   /* 1: phase-only output spectrum */
  int filter_count; /* bands; band 0 is the broad band */
  const float *responses; /* host [filter_count][nfft_filter/2+1] |H|^2 */
  const float *freq_lows; /* host [filter_count] low corners */
} fxc_s2s_config;

typedef struct fxc_s2s_plan fxc_s2s_plan;

int fxc_s2s_plan_create(fxc_s2s_plan **plan, const fxc_s2s_config *cfg);
int fxc_s2s_plan_destroy(fxc_s2s_plan *plan);
/* Derived sizes: nspec, time taper length, RUN-ABS and whitening windows. */
int fxc_s2s_plan_info(const fxc_s2s_plan *plan, int *nspec, int *taper_size, int *runabs_win,
                      int *whiten_win);
/* d_traces: [nfiles][npts] float (nfiles = groups*num_ch, channels of a group
 * adjacent); h_steps: window indices to process (skip list applied);
 * d_out: [nsteps][nfiles][nspec] complex64, the StepPack payload order. */
int fxc_s2s_process(fxc_s2s_plan *plan, const float *d_traces, int nfiles, const int *h_steps,
                    int nsteps, float *d_out, void *stream);
/* Running-mean arithmetic (RUN-ABS weights and whitening weights): 0 = FP64
 * window sums as the reference (smoothRowsRollingKernel,
 * native/sac2spec/kernels/real_matrix.cu:155-191; the default); 1 = spec
 * mode, float window sums with Neumaier compensation (no FP64 in the running
 * means).  Mode 1 needs the fused one-channel path (FXC_ERR_UNSUPPORTED
 * otherwise). */
int fxc_s2s_set_sums(fxc_s2s_plan *plan, int mode);

/* --------------------------------------------------------------- stacks --
 * Phase-weighted stacks of cross-correlations of length n (2H+1).  Groups of
 * pair p are rows h_offsets[p] .. h_offsets[p+1]-1 of d_prestack [groups][n];
 * d_weights [groups] = n_g/T; d_linear [npairs][n]; d_out [npairs][n].
 * Every n from 2 to 8192 (PWS) / 6144 (TF-PWS) is served: n with prime
 * factors in {2,3,5,7,23,29} by the direct transform, other n by Bluestein's
 * convolution.  A longer n is FXC_ERR_ARG ("too long"), raised before any
 * launch.
 * fxc_pws: prestack rows are group MEANS; out = lin * |sum_g w_g phi_g|
 * (compute_pws_host_batch, native/pws/worker/gpu.cu:87-282).
 * fxc_tfpws: prestack rows are group SUMS; Stockwell-domain PWS with
 * Gaussian scale 0.1, power 1 and the optional band blend
 * (compute_tfpws_from_prestack, native/tfpws/core/tfpws_compute.cu:19-366). */
int fxc_pws(const float *d_prestack, const float *d_weights, const long long *h_offsets,
            const float *d_linear, int npairs, int n, float *d_out, void *stream);
int fxc_tfpws(const float *d_prestack, const float *d_weights, const long long *h_offsets,
              const float *d_linear, int npairs, int n, int band_limited, float delta, float fmin,
              float fmax, float taper_hz, float *d_out, void *stream);
/* Spec mode of both stacks (no reference counterpart; the reference is
 * power 1 over every voice, native/pws/kernels/cuda.pws.cu:79-95 and
 * native/tfpws/core/tfpws_compute.cu:224).  power: the exponent of |W|
 * (> 0; 1 = the reference).  nvoices (TF-PWS only): 0 or >= n/2+1 = every
 * voice; else the groups are transformed only at the voice bins
 * v_j = floor(j (n/2) / (nvoices-1)), j < nvoices, and each bin's weight is
 * interpolated linearly between its two voices.  A null spec is the
 * reference. */
typedef struct fxc_stack_spec {
  float power;
  int nvoices;
} fxc_stack_spec;
int fxc_pws_spec(const float *d_prestack, const float *d_weights, const long long *h_offsets,
                 const float *d_linear, int npairs, int n, const fxc_stack_spec *spec,
                 float *d_out, void *stream);
int fxc_tfpws_spec(const float *d_prestack, const float *d_weights, const long long *h_offsets,
                   const float *d_linear, int npairs, int n, int band_limited, float delta,
                   float fmin, float fmax, float taper_hz, const fxc_stack_spec *spec,
                   float *d_out, void *stream);

/* PWS without device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops).  Same arguments and result as fxc_pws (power 1); its
 * workspace (offsets and the groups' analytic spectra, 8 (npairs+1) +
 * 8 n groups bytes) lives in a fixed 64 MiB device array per GPU, held until
 * the process ends; larger calls are FXC_ERR_ARG.  fxc_stack_prepare(n)
 * must run first on the device (once per device and length): it builds
 * the length-n FFT plans and resolves the array's device address (it
 * allocates; fxc_pws_noalloc then calls no hipMalloc/hipFree).  Calls on
 * one GPU share the array: calls that overlap on one GPU race. */
int fxc_stack_prepare(int n);
int fxc_pws_noalloc(const float *d_prestack, const float *d_weights, const long long *h_offsets,
                    const float *d_linear, int npairs, int n, float *d_out, void *stream);

/* Linear stack: d_out[p][j] = (float)(sum of (double)d_records[r][j] over
 * rows r = h_offsets[p] .. h_offsets[p+1]-1, divided in double by their
 * count).  Rows of a pair are summed in row order, so the result equals the
 * float64 running sum and division of _linear_stack_group
 * (fastxc/operators/stacking/linear_stack.py:228-247) bit for bit.
 * d_records [h_offsets[npairs]][n] float; npairs <= 65535 per call. */
int fxc_linear_stack(const float *d_records, const long long *h_offsets, int npairs, int n,
                     float *d_out, void *stream);

/* ------------------------------------------------------------- rotation --
 * ENZ -> RTZ rotation of stacked pair correlations; replaces
 * rotate_enz_to_rtz / build_rotate_matrix
 * (fastxc/operators/rotation/python_rotate.py:34-104), applied per pair by
 * rotate_sourcepack_stack (:264-390).  d_enz / d_rtz [npairs][9][n] float,
 * components in the order EE EN EZ NE NN NZ ZE ZN ZZ -> RR RT RZ TR TT TZ ZR
 * ZT ZZ; d_trig [npairs][4] float (sin az, cos az, sin baz, cos baz, each the
 * float sine / cosine of the float radian angle), 16-byte aligned.  Each
 * matrix entry is the float product the reference forms; each output is a
 * fused multiply-add chain in input order.  d_rtz may equal d_enz.
 * Asynchronous on the stream. */
int fxc_rotate_rtz(const float *d_enz, const float *d_trig, long long npairs, int n,
                   float *d_rtz, void *stream);

/* ------------------------------------------------------------- planning --
 * Station-pair planning on the host (no GPU): replaces build_valid_paths /
 * _build_path_record_if_valid / azimuth_in_ranges / distkm_az_baz_rudoe
 * (fastxc/inventory/planner.py:294-312, 564-617, 833-1000).  Nodes are in id
 * order (id = index + 1); lat/lon in degrees, has_geo 0 where either is
 * unknown; group ids compare equal for equal group names.  Pairs (i, j),
 * j >= i (j > i without autocorrelation, j == i only with autocorr_only),
 * pass the group rule (group_mode 0 all, 1 same group, 2 cross group; never
 * applied to autocorrelations), then the distance [min_km, max_km] and the
 * azimuth ranges (az_ranges [nranges][2] degrees, either the azimuth or the
 * back azimuth inside).  The geometry is float64 libm arithmetic in the
 * reference's order (bit-identical values).  Threads: nthreads (0 = all
 * cores).  Results stay in an opaque set until fetched: npaths pairs in
 * (src, rec) order, nchecks pairs examined. */
int fxc_plan_paths(int nnodes, const double *lat, const double *lon,
                   const unsigned char *has_geo, const int *group, int allow_autocorr,
                   int autocorr_only, int group_mode, int use_distance, double min_km,
                   double max_km, int use_azimuth, const double *az_ranges, int nranges,
                   int nthreads, void **out_set, long long *npaths, long long *nchecks);
/* src/rec [npaths] node indices, has_geo [npaths], geo [npaths][4] (great
 * circle degrees, azimuth, back azimuth, distance km; 0 without geometry). */
int fxc_plan_paths_fetch(void *set, int *src, int *rec, unsigned char *has_geo, double *geo);
/* allowed_paths.tsv and allowed_path_ids.txt of the set as write_path_plan
 * (fastxc/inventory/planner.py:351-395) writes them: csv.writer rows (tab,
 * minimal quoting, "\r\n"), ids "%08d" / "%04d", node_text[i] the node's
 * group.network.station.location, numbers "%.7g", empty without geometry. */
int fxc_plan_paths_write(void *set, int nnodes, const int *node_id, const char *const *node_text,
                         const char *tsv_path, const char *ids_path, int nthreads);
int fxc_plan_paths_free(void *set);

/* ------------------------------------------------------------ streaming --
 * Serving loop over the stages above, keeping per-pair stacks in HBM.  The
 * reference has no counterpart (it stacks offline: the linear stack of
 * fastxc/operators/stacking/linear_stack.py:228-247 and ncf_pws / ncf_tfpws
 * over stored correlations); the per-segment work is fxc_s2s_process and
 * fxc_xc_correlate, and the state update is the linear stack's float64
 * running sum.
 *
 * A segment is the traces [n][cfg.npts] (host float) of n channels (ids
 * 0..nchan-1, in groups of cfg.num_ch); every channel pair (a, b), a < b by
 * id, is correlated and added to that pair's state: a float64 accumulator of
 * the 2H+1 lags, a record count, and (history > 0) the last `history`
 * correlations.  `lanes` segments are processed concurrently, each lane with
 * its own stream and workspaces for up to max_seg_chan channels, allocated
 * at creation; state updates are applied one segment at a time.  At most
 * max_pairs pairs have state; a segment with more new pairs than free slots
 * is refused (FXC_ERR_ALLOC from fxc_stream_wait).  history must be 0 or at
 * least `lanes`. */
typedef struct fxc_stream fxc_stream;
int fxc_stream_create(fxc_stream **stream, const fxc_s2s_config *cfg, int nchan, int half_cc,
                      int precision, int history, long long max_pairs, int max_seg_chan,
                      int lanes);
/* Queues a segment; h_traces must stay valid until its wait returns.  The
 * latency fxc_stream_wait reports runs from this call to the state update. */
int fxc_stream_submit(fxc_stream *stream, const int *h_chan, int n, const float *h_traces,
                      long long *ticket);
int fxc_stream_wait(fxc_stream *stream, long long ticket, double *latency_ms);
int fxc_stream_info(fxc_stream *stream, long long *pairs_used, long long *segments_done,
                    long long *state_bytes, long long *lane_bytes);
/* After every submitted segment is applied: kind 0 the linear stack
 * (float)(acc / count), 1 PWS and 2 TF-PWS of the history in submission
 * order with sub-stacks of `substack` records (band arguments as fxc_tfpws).
 * Pairs are channel ids src < rec; d_out [npairs][2H+1]; h_count (optional)
 * receives each pair's record count. */
int fxc_stream_read(fxc_stream *stream, int kind, const int *h_src, const int *h_rec,
                    long long npairs, int substack, int band_limited, float fmin, float fmax,
                    float taper_hz, float *d_out, long long *h_count, void *hip_stream);
/* Running PWS state (call once, before the first segment): every pair then
 * keeps sum_g n_g phi_g over its closed sub-stacks of `substack` records
 * and the sum of its open sub-stack (12 B per lag per pair of max_pairs;
 * 8 B with substack 1, whose sub-stacks close as they arrive),
 * so fxc_stream_read kind 1 returns the PWS of ALL its records in
 * submission order, equal to fxc_pws over them with sub-stacks of
 * `substack` (the reference's PWS, native/pws/io/sourcepack_io.cu:385-478,
 * native/pws/kernels/cuda.pws.cu:28-95; lin is the float64 linear stack). */
int fxc_stream_set_pws(fxc_stream *stream, int substack);
/* fxc_stream_read with the stacks' spec mode (power; voices for TF-PWS). */
int fxc_stream_read_spec(fxc_stream *stream, int kind, const int *h_src, const int *h_rec,
                         long long npairs, int substack, int band_limited, float fmin, float fmax,
                         float taper_hz, const fxc_stack_spec *spec, float *d_out,
                         long long *h_count, void *hip_stream);
int fxc_stream_destroy(fxc_stream *stream);

#ifdef __cplusplus
}
#endif

#endif /* FASTXC_AMD_H */
