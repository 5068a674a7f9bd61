/* C-ABI of the MI355X Lucas-Lehmer engine (libcudalucas_amd.so).
 *
 * This is the drop-in boundary for the LL iteration path of
 * tdulcet/CUDALucas v2.06.  The reference has no library API: its path sits
 * behind internal functions and process globals.  Each entry point below
 * names the reference function (CUDALucas.cu:line) whose role it takes.
 * Plain C types only; every call returns an int status (LL_OK or a negative
 * LL_ERR_*) instead of exit()ing (the reference exits from
 * cuda_safecalls.h:43-48), and ll_last_error() explains the last failure.
 *
 * Digits crossing the boundary are int32 in natural order, value
 * sum_j d_j 2^ceil(p j / n) mod 2^p - 1 of the *shifted* residue
 * s * 2^offset, exactly as the reference's x_int / g_xint arrays
 * (CUDALucas.cu:1254-1342).
 *
 * Concurrency: a context (or batch) runs its kernels on its own stream(s).
 * The carry sweep of the two-sweep path waits on other workgroups of the
 * same launch; two such launches side by side on one GPU could hold each
 * other's workgroup slots.  Every call that runs carry sweeps (ll_iterate,
 * ll_iterate_noalloc, ll_profile_kernels, ll_batch_step) therefore holds the
 * GPU's lock from its first launch to its synchronisation: a mutex for the
 * threads of the process and an flock on
 * "$TMPDIR/cudalucas_amd-gpu-<PCI bus id>.lock" (/tmp without TMPDIR) for
 * other processes, so contexts in different threads or processes can
 * iterate at the same time and their calls interleave.  One context must
 * not be used from two threads at once.  A wait that does not end is
 * reported as a failed iteration (error 3.0), never a hang.
 */
#ifndef LL_API_H
#define LL_API_H
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LL_OK 0
#define LL_ERR_ARG (-1)         /* invalid argument */
#define LL_ERR_UNSUPPORTED (-2) /* FFT length / schedule not available */
#define LL_ERR_HIP (-3)         /* HIP runtime or device error */
#define LL_ERR_STATE (-4)       /* call out of order (e.g. iterate before load) */
#define LL_ERR_ALLOC (-5)       /* device or host allocation failed */
#define LL_ERR_ROUNDOFF (-6)    /* unload rounding error: state corrupt */

typedef struct ll_ctx ll_ctx;

typedef struct ll_options {
  int32_t device;      /* HIP device index (-d, CUDALucas.cu:3655-3665) */
  int32_t threads[2];  /* -threads t1 t2 (CUDALucas.cu:812-945); 0 = default.
                          Mapped to the carry sweep's lines per workgroup
                          (t2/32, clamped to {2,4,8}). */
  uint32_t width;      /* schedule override: line length W (0 = auto) */
  uint32_t height;     /* schedule override: column length H (0 = auto) */
  uint32_t lines;      /* schedule override: lines per carry workgroup R */
  uint32_t flags;      /* LL_OPT_* */
} ll_options;

#define LL_OPT_FORCE_SWEEPS 1u /* use the two-sweep path even if LDS-resident fits */
#define LL_OPT_NO_GRAPH 2u     /* launch kernels one by one (no hipGraph replay) */
#define LL_OPT_GENERIC 4u      /* two sweeps on the runtime-plan (generic) kernels, which
                                  every length without a specialised schedule uses */

typedef struct ll_info {
  uint32_t p, n;
  uint32_t path; /* 0 = LDS-resident single workgroup, 1 = two HBM sweeps,
                    2 = two HBM sweeps on the generic (runtime-plan) kernels */
  uint32_t width, height, lines;
  void *stream;  /* the hipStream_t all work of this context runs on */
  uint64_t iterations; /* iterations run since the last load */
} ll_info;

/* Replaces init_threads + alloc_gpu_mem + write_gpu_data + cufftPlan1d
 * (CUDALucas.cu:812, 1096, 1125, 1239): builds tables and buffers for
 * exponent p at FFT length n (words). */
int ll_create(uint32_t p, uint32_t n, const ll_options *opts, ll_ctx **out);

/* Replaces close_lucas / free_gpu (CUDALucas.cu:1400-1419). */
void ll_destroy(ll_ctx *ctx);

/* Replaces init_x + apply_weights (CUDALucas.cu:216-249, 1254-1278):
 * loads n int32 digits of the shifted residue with shift `offset`.
 * Digits should be balanced (ll_balance_digits) for full FFT accuracy. */
int ll_load_digits(ll_ctx *ctx, const int32_t *digits, uint32_t offset);

/* Replaces `count` calls of lucas_square (CUDALucas.cu:1869-1986): each
 * iteration doubles the offset mod p, squares, subtracts 2 at bit
 * offset+1, and records its exact maximum |x - rint(x)| (the reference
 * samples it every ErrorIterations, CUDALucas.cu:1895-1900).
 * per_iter_err (may be NULL) receives `count` floats; max_err (may be NULL)
 * the maximum over them.  A value >= 1.0 flags a carry that overflowed the
 * 32-bit carry shuttle: the iteration must be treated as failed. */
int ll_iterate(ll_ctx *ctx, uint32_t count, float *per_iter_err, float *max_err);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory while they run (serving loops, harnesses that forbid hipMalloc):
 * ll_iterate with the carry shuttle and the per-iteration error slots in a
 * file-scope device array of LL_NOALLOC_SCRATCH_BYTES per GPU instead of
 * the context's own buffers.  It calls no hipMalloc/hipFree and replays no
 * graph (direct kernel launches; results identical to ll_iterate).  Limit:
 * the schedule's shuttle, (H / R) * W * 8 bytes, must fit in
 * LL_NOALLOC_SCRATCH_BYTES - 64 KiB (LL_ERR_ARG otherwise; e.g. 2 MiB at
 * 4M words).  Every call on a GPU shares that GPU's array, which keeps the
 * last call's contents: each call clears the shuttle first, and calls that
 * overlap on one GPU (another thread, another context) race. */
#define LL_NOALLOC_SCRATCH_BYTES (16u << 20)
int ll_iterate_noalloc(ll_ctx *ctx, uint32_t count, float *per_iter_err, float *max_err);

/* Replaces cudaMemcpy(x_int, g_xint) (CUDALucas.cu:2773, 2910, 2963, 3112,
 * 3155): n int32 digits of the current shifted residue and its offset. */
int ll_read_digits(ll_ctx *ctx, int32_t *digits, uint32_t *offset);

/* Replaces reset_err (CUDALucas.cu:1627-1631): clears the running maximum
 * error reported by ll_error_max. */
int ll_reset_error(ll_ctx *ctx);
float ll_error_max(const ll_ctx *ctx);

int ll_get_info(const ll_ctx *ctx, ll_info *info);

/* GPU time (ms, hipEvents on the context's stream) of the kernels launched
 * by the last ll_iterate call, excluding host readback. */
double ll_last_iterate_ms(const ll_ctx *ctx);

/* Per-kernel timing, the analog of kernel_percentages
 * (CUDALucas.cu:2002-2140): runs `count` iterations, each kernel launched
 * with a pair of hipEvents that take its own start and end
 * (hipExtLaunchKernel, so the dispatch gaps between launches are not
 * counted), and returns the average duration in ms:
 * kernel_ms[0] = square sweep (or one whole iteration of the LDS-resident
 * kernel), kernel_ms[1] = carry sweep (0 for the resident path).  Per-
 * iteration errors are not collected. */
int ll_profile_kernels(ll_ctx *ctx, uint32_t count, float *kernel_ms);
const char *ll_last_error(const ll_ctx *ctx);

/* Measurement helper, no reference counterpart: average ms of a plain
 * device-to-device streaming copy of `bytes` (one read + one write), the
 * calibration for HBM byte counters and achievable bandwidth. */
int ll_bench_copy(int32_t device, uint64_t bytes, uint32_t reps, float *ms);

/* Test hook, no reference counterpart (the reference can only wait for a
 * real failure, CUDALucas.cu:2938-2950, 3043): makes the iteration that is
 * `after` iterations ahead report `value` as its error (kind 1), or makes
 * the next ll_iterate fail with LL_ERR_HIP as a device error would (kind 2).
 * kind 0 clears. */
int ll_debug_inject(ll_ctx *ctx, uint32_t kind, uint64_t after, float value);

/* ---- batched engine (no reference counterpart: CUDALucas runs one test
 * per process) --------------------------------------------------------- */
typedef struct ll_batch ll_batch;
typedef struct ll_batch_test {
  uint32_t p, n;       /* exponent and FFT length (words) */
  uint32_t offset;     /* shift offset of s_0 = 4 */
  uint32_t iterations; /* iterations to run (p - 2 for a full test) */
} ll_batch_test;
/* Builds every test's tables and state (s_0 = 4 at its offset), groups the
 * tests by FFT length and schedule. flags: LL_OPT_FORCE_SWEEPS. */
int ll_batch_create(int32_t device, uint32_t count, const ll_batch_test *tests, uint32_t flags,
                    ll_batch **out);
void ll_batch_destroy(ll_batch *batch);
/* Advances every unfinished test by up to `iterations` (<= 1024) LL
 * iterations; step_max_err (may be NULL) receives, per test, the maximum
 * roundoff of the iterations it ran in this step. */
int ll_batch_step(ll_batch *batch, uint32_t iterations, float *step_max_err);
int ll_batch_status(const ll_batch *batch, uint32_t index, uint64_t *iterations_done,
                    uint32_t *remaining, float *max_err);
/* Current digits of test `index` (as ll_read_digits). */
int ll_batch_read_digits(ll_batch *batch, uint32_t index, int32_t *digits, uint32_t *offset);
double ll_batch_last_step_ms(const ll_batch *batch);
/* Per-iteration roundoff of test `index` over the last ll_batch_step (one
 * value per iteration it ran, in order).  Returns the count, or LL_ERR_ARG
 * if `capacity` is too small; errs may be NULL to query the count. */
int ll_batch_step_errors(const ll_batch *batch, uint32_t index, float *errs, uint32_t capacity);
int ll_batch_groups(const ll_batch *batch);
const char *ll_batch_last_error(const ll_batch *batch);

/* ---- roundoff surrogate for FFT-length selection (no reference
 * counterpart: the reference picks lengths from fixed power laws,
 * CUDALucas.cu:1027-1068, and corrects with roundoff retries,
 * CUDALucas.cu:2600-2640).  An FP32 MLP, tanh hidden layers and one linear
 * output, trained with mean squared error and Adam, all on the GPU.
 * Parameters are one vector: per layer l, W_l [sizes[l+1] x sizes[l]]
 * row-major then b_l [sizes[l+1]]. Widths <= 128; sizes[layers] must be 1. */
typedef struct ll_mlp ll_mlp;
int ll_mlp_create(int32_t device, uint32_t layers, const uint32_t *sizes, ll_mlp **out);
void ll_mlp_destroy(ll_mlp *mlp);
const char *ll_mlp_last_error(const ll_mlp *mlp);
uint32_t ll_mlp_param_count(const ll_mlp *mlp);
/* Replaces the parameters and clears the Adam moments. */
int ll_mlp_set_params(ll_mlp *mlp, const float *params);
int ll_mlp_get_params(ll_mlp *mlp, float *params);
/* y[batch] = f(x[batch][sizes[0]]) */
int ll_mlp_forward(ll_mlp *mlp, const float *x, uint32_t batch, float *y);
/* Loss = mean (f(x) - t)^2 over the batch; grad (may be NULL) receives
 * dLoss/dparams, loss (may be NULL) the loss. Parameters are unchanged. */
int ll_mlp_gradient(ll_mlp *mlp, const float *x, const float *t, uint32_t batch, float *grad,
                    float *loss);
/* `epochs` passes of minibatch Adam (beta 0.9 / 0.999, eps 1e-8) over
 * `count` samples in seeded shuffled order; a trailing partial batch is
 * dropped. epoch_loss (may be NULL) receives each epoch's mean batch loss. */
int ll_mlp_train(ll_mlp *mlp, const float *x, const float *t, uint32_t count, uint32_t batch,
                 uint32_t epochs, float lr, uint64_t seed, float *epoch_loss);

/* The selector: features of (p, n) -- log2 p, log2 n, bits per word,
 * exponents of 3, 5 and 7 in n, 1 if n runs LDS-resident -- feed the MLP,
 * whose output is log2 of the maximum roundoff over a window of iterations
 * (after normalisation; both affine maps are stored in the model file).
 * The prediction is extrapolated to the p - 2 iterations of a test by the
 * sqrt(ln(#outputs)) growth of a Gaussian maximum and raised by the model's
 * calibrated margin. */
#define LL_SELECTOR_FEATURES 7
typedef struct ll_selector ll_selector;
/* Writes LL_SELECTOR_FEATURES raw features; returns their count or < 0. */
int ll_selector_features(uint32_t p, uint32_t n, float *features);
/* Lengths the selector considers for p: supported multiples of 1K words in
 * [0.62, 1.25] x the reference policy's length, at most 28 bits per word.
 * Returns the count (ns may be NULL to query it). */
int ll_selector_candidates(uint32_t p, uint32_t *ns, uint32_t capacity);
/* Loads a text model ("cudalucas_amd roundoff-mlp 1") onto `device`. */
int ll_selector_load(int32_t device, const char *path, ll_selector **out);
void ll_selector_destroy(ll_selector *sel);
const char *ll_selector_last_error(const ll_selector *sel);
/* Measured ms/iteration per length (the fft bench table, or per-test cost
 * in the batched engine); with a table the selector takes the fastest of the
 * lengths predicted safe in [0.62, 1.25] x the reference policy's and the
 * longer ones up to 2 x (safe without a prediction); without one, the
 * shortest length predicted safe no longer than the reference policy's. */
int ll_selector_set_times(ll_selector *sel, uint32_t count, const uint32_t *ns, const float *ms);
/* log2 of the predicted full-test maximum roundoff at each length: the
 * network's window prediction + growth_scale x the Gaussian growth term +
 * margin (growth_scale and margin are calibrated on held-out traces). */
int ll_selector_predict(ll_selector *sel, uint32_t p, uint32_t count, const uint32_t *ns,
                        float *log2_err);
/* Replaces choose_fft_length (CUDALucas.cu:1027-1068) when a model is
 * loaded.  above = 0: the first choice (see ll_selector_set_times; the
 * reference policy's length if none is predicted safe).  above = n: the
 * fallback after length n failed the round-off test (CUDALucas.cu:2589-2648):
 * the best acceptable length longer than n.  pred_err (may be NULL)
 * receives the predicted full-test maximum. */
int ll_selector_choose(ll_selector *sel, uint32_t p, uint32_t above, float *pred_err);

/* ---- host codec (CUDALucas.cu:1280-1342, 1465-1614) ------------------ */
int ll_digit_width(uint32_t p, uint32_t n, uint32_t j);
/* unpack_bits (CUDALucas.cu:1285-1311) */
void ll_unpack_bits(uint32_t p, uint32_t n, const uint32_t *packed, int32_t *digits);
/* balance_digits (CUDALucas.cu:1313-1342) */
void ll_balance_digits(uint32_t p, uint32_t n, int32_t *digits);
/* standardize_digits (CUDALucas.cu:1465-1489), exact for any int32 digits:
 * full cyclic carry to [0, 2^w) digits of the value mod 2^p-1.
 * Returns 1 if the value is 0. */
int ll_standardize_digits(uint32_t p, uint32_t n, int32_t *digits);
/* pack_bits (CUDALucas.cu:1594-1614): standard digits -> ceil(p/32) words */
void ll_pack_bits(uint32_t p, uint32_t n, const int32_t *digits, uint32_t *packed);
/* Unshifted residue (value * 2^-offset mod 2^p-1) as ceil(p/32) words. */
int ll_residue_words(uint32_t p, uint32_t n, const int32_t *digits, uint32_t offset,
                     uint32_t *words);
/* find_residue (CUDALucas.cu:1491-1534): low 64 bits of the unshifted
 * residue. */
uint64_t ll_res64(uint32_t p, uint32_t n, const int32_t *digits, uint32_t offset);
/* init_x for a new test (CUDALucas.cu:1261-1267): 4 * 2^offset. */
void ll_initial_digits(uint32_t p, uint32_t n, uint32_t offset, int32_t *digits);
/* checkpoint_checksum (CUDALucas.cu:1712-1733) */
uint32_t ll_checkpoint_crc(const void *bytes, size_t len);

/* ---- checkpoint files (CUDALucas.cu:1616-1625, 1707-1861) ------------- */
/* Layout: uint32 words[ceil(p/32) + 10]; [0, end) the packed shifted residue
 * (standardized digits, pack_bits), then p, n, next iteration, offset,
 * total_us >> 15, time_adj_us >> 15, iter_adj, 0, magic 0, CRC of the
 * preceding 4 (end + 9) bytes. */
typedef struct ll_checkpoint {
  uint32_t p, n, next_iter, offset;
  uint64_t total_us, time_adj_us;
  uint32_t iter_adj;
} ll_checkpoint;
size_t ll_checkpoint_words(uint32_t p);
/* set_checkpoint_data + standardize_digits + pack_bits + CRC */
int ll_checkpoint_encode(const ll_checkpoint *meta, const int32_t *digits, uint32_t *words);
/* Validates size fields and CRC; fills meta. Returns LL_OK or LL_ERR_ARG. */
int ll_checkpoint_decode(uint32_t p, const uint32_t *words, ll_checkpoint *meta);
/* Digits at length n (unpack_bits + balance_digits) of a decoded checkpoint. */
int ll_checkpoint_digits(uint32_t p, uint32_t n, const uint32_t *words, int32_t *digits);
/* write_checkpoint: rotates <dir>/c<p> to <dir>/t<p>, writes <dir>/c<p>;
 * with save_dir non-NULL also <save_dir>/s<p>.<iter>.<res64 hex>. */
int ll_checkpoint_write(const char *dir, uint32_t p, const uint32_t *words, const char *save_dir,
                        uint64_t res64);
/* read_checkpoint: c<p>, falling back to t<p>; each must pass its CRC (the
 * reference checks the backup with a plain sum, CUDALucas.cu:1791).
 * Returns 0 = c file, 1 = t file, LL_ERR_ARG = neither usable. */
int ll_checkpoint_read(const char *dir, uint32_t p, uint32_t *words);
/* rm_checkpoint (CUDALucas.cu:1851-1861) */
void ll_checkpoint_remove(const char *dir, uint32_t p);

/* ---- FFT length policy (CUDALucas.cu:947-1094) ----------------------- */
int ll_fft_length_count(void);
int ll_fft_length_at(int index); /* words, from the reference's default list */
/* 1 if this build has kernels for length n */
int ll_fft_supported(uint32_t n);
/* The schedule ll_create would use for length n under opts (may be NULL):
 * fills info->n, path, width, height, lines (resident: width = n/2,
 * height = lines = 1). LL_ERR_UNSUPPORTED if there is none. The analog of
 * the reference's thread/plan tables (CUDALucas.cu:812-945). */
int ll_fft_schedule(uint32_t n, const ll_options *opts, ll_info *info);

/* ---- schedule tuner (the analog of -threadbench, CUDALucas.cu:2143-2276,
 * whose "<device> threads.txt" init_threads reads back, :812-945) -------- */
typedef struct ll_schedule {
  uint32_t n;                    /* FFT length (words) */
  uint32_t path;                 /* 0 = LDS-resident, 1 = two sweeps */
  uint32_t width, height, lines; /* W, H, R (resident: n/2, 1, 1) */
  float ms;                      /* measured ms per iteration (< 0: failed) */
} ll_schedule;
/* Every schedule this build has for length n; returns the count (out may
 * be NULL to query it). */
int ll_fft_schedules(uint32_t n, ll_schedule *out, uint32_t capacity);
/* Times every schedule of length n over `iters` hipGraph-replayed
 * iterations (after a warm-up) at exponent p (0: about 17 bits per word);
 * out receives them fastest first. Returns the count. */
int ll_tune_length(int32_t device, uint32_t p, uint32_t n, uint32_t iters, ll_schedule *out,
                   uint32_t capacity);
/* Installs preferred schedules (e.g. read back from a tuning file): ll_create
 * and ll_batch_create use the entry for n unless the options name a schedule
 * (width/height/lines, -threads, LL_OPT_FORCE_SWEEPS). count = 0 clears. */
int ll_schedule_table_set(uint32_t count, const ll_schedule *table);

/* Installs the FFT-length list that ll_fft_length_at / ll_choose_fft_length
 * use (ascending, words), as init_ffts merges "<device> fft.txt" into the
 * defaults (CUDALucas.cu:947-1025). count = 0 restores the 162 defaults. */
int ll_fft_length_list_set(uint32_t count, const uint32_t *lengths);

/* Device properties for -info and the tuning files (init_device,
 * CUDALucas.cu:737-810). */
typedef struct ll_device_info {
  char name[256];
  char arch[64];
  int32_t compute_units, clock_mhz, mem_clock_mhz, mem_bus_bits;
  uint64_t global_mem;
  int32_t l2_bytes;
  uint32_t lds_per_block;
  int32_t max_threads_per_block, wavefront, pci_bus, pci_device;
  int32_t runtime_version, driver_version;
} ll_device_info;
int ll_get_device_info(int32_t device, ll_device_info *out);
/* choose_fft_length (CUDALucas.cu:1027-1068): *index < 0 selects from the
 * list by the lb/ub power laws; *index >= count is a manual length.  Lengths
 * without kernels in this build are skipped upward. Returns n or < 0. */
int ll_choose_fft_length(uint32_t p, int *index);
void ll_fft_bounds(uint32_t p, int32_t *lb, int32_t *ub);
/* fft_from_str (CUDALucas.cu:1070-1094): "4096K", "4M", "4194304"; -1 on error */
int ll_fft_from_str(const char *s);

/* ---- worktodo file (parse.c:155-678) ---------------------------------- */
/* One parsed work line (struct ASSIGNMENT, parse.c:230-236). */
#define LL_WORK_TEXT 192
typedef struct ll_assignment {
  uint32_t exponent;           /* the largest plain number on the line */
  uint32_t fft_length;         /* words, from a "<k>K" or "<m>M" field; 0 = none */
  char key[LL_WORK_TEXT];      /* the field holding a letter a-f (hex key, N/A); "" = none */
  char comment[LL_WORK_TEXT];  /* the text after a #, // or \\ delimiter */
} ll_assignment;
/* Line status (enum PARSE_WARNINGS, parse.c:218-228). */
#define LL_WORK_OK 0
#define LL_WORK_EOF 1
#define LL_WORK_LONG_LINE 2      /* 191 characters or more */
#define LL_WORK_NO_TEST_EQUAL 3  /* neither "Test=" nor "DoubleCheck=" */
#define LL_WORK_INVALID_FORMAT 4 /* more than 5 fields, a field without a number */
#define LL_WORK_INVALID_DATA 5   /* out of range: exponent < 100, length < 4096 */
#define LL_WORK_BLANK 6
#define LL_WORK_COMMENT 7        /* a line starting with #, // or \\ */
/* parse_worktodo_line (parse.c:241-484) on one line; a trailing newline is
 * ignored. Fills out (zeroed first) and returns an LL_WORK_* status
 * (LL_ERR_ARG if out is NULL). */
int ll_work_parse_line(const char *line, ll_assignment *out);
/* isprime (parse.c:155-174): 1 if n is prime. */
int ll_is_prime(uint32_t n);
/* valid_assignment (parse.c:176-203): 1 = valid; 0 = exponent < 7500 or a
 * length that is not a multiple of 1024; -1 = exponent not prime. Prints the
 * reference's warning on stderr. */
int ll_valid_assignment(uint32_t p, uint32_t fft_length);
/* get_next_assignment (parse.c:486-554): the first line that parses and is
 * valid; each other non-blank, non-comment line before it is reported on
 * stderr and skipped. Returns 0, 1 = cannot open the file, 2 = none found
 * (LL_ERR_ARG for a NULL argument). */
int ll_work_next(const char *file, ll_assignment *out);
/* clear_assignment (parse.c:556-678): rewrites the file without the first
 * assignment for exponent p.  The comment, blank and unparsable lines just
 * above it go with it: back to the previous assignment line, or, with none
 * above, back to the file's first blank line (lines before that stay).
 * Returns 0, 3 = cannot open the file, 4 = cannot create the temporary
 * "<file>.tmp", 5 = no such assignment (file unchanged), 6 = cannot replace
 * the file (LL_ERR_ARG for a NULL file). */
int ll_work_clear(const char *file, uint32_t p);

/* ---- the report line (process_output + encode_output_options,
 * CUDALucas.cu:1633-1701, 3189-3372) ------------------------------------ */
typedef struct ll_report {
  uint32_t p, n, iteration, offset;
  uint64_t residue;              /* Res64 */
  float max_err;                 /* maximum roundoff since the last report */
  uint32_t interval_iterations;  /* iterations since the last report */
  uint64_t interval_us;          /* their wall time, microseconds */
  uint64_t eta_s;                /* estimated seconds to the end of the test */
  int64_t now;                   /* time_t of the report, for the date and clock */
  const char *program;           /* program name and version */
} ll_report;
/* Formats one report line by the ini's OutputString: %a %M %D %h %m %s %y %z
 * (weekday, month, day, hour, :minutes, :seconds, year, zone; local time),
 * %p %P %f %F %i %o %r %R %N %% (exponent, M<p>, length in K, length,
 * iteration, offset, Res64 lower/upper case, program, %), %C %T (interval
 * h:mm:ss, ETA d:hh:mm:ss, right-aligned in 12), and %x %q %c %Q %d followed by
 * a digit k = 0..5 (roundoff, ms/iteration, interval seconds, iterations/s,
 * percent done, printed with 5 decimals and cut to k).  Other text is
 * copied.  Writes at most cap bytes (NUL-terminated) and returns the length
 * of the whole line, as snprintf does (LL_ERR_ARG for a NULL format or r;
 * out may be NULL with cap 0 to ask for the length). */
int ll_format_report(const char *format, const ll_report *r, char *out, size_t cap);

/* ---- locked files (fopen_and_lock / unlock_and_fclose, parse.c:775-894) - */
/* Opens path with a lock: "<path>.lck", created exclusively, is the lock,
 * so this build and the reference exclude each other.  A lock held
 * elsewhere is waited for ("<path> is locked, waiting ..." once, sleeps
 * growing to 1 s); wait_ms = 0 waits as long as it takes, as the reference
 * does, otherwise NULL with errno = ETIMEDOUT after about wait_ms.  Where the
 * lock file cannot be made at all (not an existing lock) the file is opened
 * without one, as the reference does.  NULL if fopen fails. */
FILE *ll_fopen_locked(const char *path, const char *mode, uint32_t wait_ms);
/* Closes a file from ll_fopen_locked and removes its lock; fclose's result. */
int ll_fclose_locked(FILE *f);

#ifdef __cplusplus
}
#endif
#endif
