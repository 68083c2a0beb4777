/* C ABI of libstl_bridge.so, the MI355X ScatCov engine.
 *
 * Drop-in for the reference's stl-bridge C ABI (crates/stl-bridge/src/lib.rs)
 * as bound by python/stl_ensemble/stl_port/native.py:55-116: same symbol
 * names, argument meanings, shapes and failure values.  Maps are float64
 * NESTED [batch, 12 nside^2]; values and cotangents are interleaved (re, im)
 * float64 pairs [batch, ncoef]; the gradient is that of
 * Re sum(conj(cotangent) * values).  Every forward state carries an identity
 * (handle, generation, member, objective); stlp_vjp and stlp_restore refuse a
 * stale or foreign one.  Failures return -1 or NULL and leave a message for
 * stlp_last_error.  One caller thread per handle.
 *
 * Scope of this library: the GPU backend of the complete one-channel plan
 * (mean, var, S1..S4; kernel 5, 4 orientations, gauge cosmo, fewer = false,
 * dg(j) = j) and the cross plan of map pairs ("cross": true), whole-sky or
 * masked averages.  The other products the reference's bridge also exports
 * (patches, spectra, U-Net, training, spin 2) are
 * present as symbols that fail with "not implemented".
 */
#ifndef STL_BRIDGE_H
#define STL_BRIDGE_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct StlHandle StlHandle;
typedef struct StlSnapshot StlSnapshot;

/* Message of the last failure on the calling thread (lib.rs:257-263). */
const char* stlp_last_error(void);
/* Frees a string returned by stlp_info (lib.rs:265-271). */
void stlp_string_free(char* s);

/* New handle from a JSON configuration (lib.rs:273-351): nside (128), j (7),
 * batch (1), cross (false: maps [batch, 12 nside^2]; true: pairs
 * [batch, 2, 12 nside^2] and their cross plan, lib.rs:238-245), families
 * (all; "S1,S2" ...: the plan of those families, plan.rs:80-100), kernel (5),
 * orientations (4), block_pixels (accepted, unused:
 * the whole sky is resident), backend ("gpu"; "cpu" is refused), device (0),
 * and the keys of this library: precision ("fp64" | "mixed": FP32 bulk
 * fields and operators with FP64 sky-wide sums), timing (false), fields
 * ("auto" | "keep" | "recompute": the field store), batch_nside (16: the
 * levels n >= 2 whose pixels over the batch number at most 12 batch_nside^2
 * run batched, one launch per sweep step for all of them; 0: none).  NULL on
 * failure. */
StlHandle* stlp_new(const char* config);
/* Frees a handle (lib.rs:353-359); NULL is ignored. */
void stlp_free(StlHandle* h);
/* A handle on the level fields and operators of h (no counterpart symbol in
 * the reference: its GpuEvaluator::share, crates/stl-gpu/src/engine.rs
 * :1909-1916, several plans and batches on one engine).  JSON: batch (h's;
 * at most h's maps, pairs counting 2), cross (h's), families ("mean,var,S1"
 * ...: a plan of those families only; default all).  Its own identity,
 * masks and sums; the handles take turns on the fields: a stlp_vjp or
 * stlp_checkpoint needs the last forward on the fields to be its own handle's.
 * Either handle may be freed first.  One caller thread for all of them. */
StlHandle* stlx_share(StlHandle* h, const char* config);
/* JSON description: version, sizes, plan, dtype, device memory, timers
 * (lib.rs:361-391).  Free with stlp_string_free; NULL on failure. */
char* stlp_info(StlHandle* h);

/* Forward sweep of x (n = batch * 12 nside^2 float64, host memory): values
 * to out (2 * batch * K * ncoef, K = 1 without masks), with moments the
 * spatial variances of the terms to var (batch * K * ncoef) (lib.rs:474-530).
 * 0 or -1. */
int stlp_forward(StlHandle* h, const double* x, size_t n, uint64_t generation, uint64_t member, uint64_t objective,
                 bool moments, double* out, double* var);
/* Map cotangent out (batch * 12 nside^2, host) of the cotangents cot
 * (2 n, n = batch * K * ncoef) after the forward of the same identity
 * (lib.rs:532-581).  0 or -1. */
int stlp_vjp(StlHandle* h, const double* cot, size_t n, uint64_t generation, uint64_t member, uint64_t objective,
             double* out);

/* As stlp_forward with x a device address (float64, on the handle's device),
 * the additive device-pointer form the reference gives its patch handles
 * (stlp_patch_forward_device, lib.rs:1152-1176).  Values stay host arrays. */
int stlp_forward_device(StlHandle* h, uint64_t x_dev, size_t n, uint64_t generation, uint64_t member,
                        uint64_t objective, bool moments, double* out, double* var);
/* As stlp_vjp with the map cotangent written to a device address (float64)
 * (cf. stlp_patch_vjp_device, lib.rs:1152-1176). */
int stlp_vjp_device(StlHandle* h, const double* cot, size_t n, uint64_t generation, uint64_t member,
                    uint64_t objective, uint64_t out_dev);

/* Pixels each coefficient averages ([K, ncoef] size_t, K = max(masks, 1))
 * (lib.rs:456-472). */
int stlp_native_pixels(StlHandle* h, size_t* out);
/* Forward on maps given at every level (concatenated [batch, 12 (nside/2^d)^2]
 * over levels d), no decimation; values as stlp_forward; no vjp may follow
 * (lib.rs:583-623). */
int stlp_forward_levels(StlHandle* h, const double* x, size_t n, double* out);

/* Host copy of the current forward state and its restore (lib.rs:645-709). */
StlSnapshot* stlp_checkpoint(StlHandle* h);
void stlp_checkpoint_free(StlSnapshot* s);
size_t stlp_checkpoint_bytes(const StlSnapshot* s);
int stlp_restore(StlHandle* h, const StlSnapshot* s, uint64_t generation, uint64_t member, uint64_t objective);

/* loss = callback(x, n, grad, context) fills grad at x (lib.rs:711-712). */
typedef double (*stlp_callback)(const double* x, size_t n, double* grad, void* context);
/* L-BFGS with strong Wolfe line search of the callback from x (n, updated in
 * place); JSON options lr (1), max_iter (20), max_eval (max_iter * 5 / 4),
 * tolerance_grad (1e-12), tolerance_change (1e-15), history_size (10).
 * Returns the number of evaluations or -1 (lib.rs:714-741). */
int64_t stlp_lbfgs(double* x, size_t n, const char* options, stlp_callback callback, void* context);

/* Averages over k masks from the next forward on (lib.rs:393-422):
 * weights [k, 12 nside^2] (nonzero: in the mask), level n + 1 holding the
 * parents of the level-n mask (mask.rs:43-70); every coefficient is then the
 * mean of its terms over the mask at its native level, and values, variances
 * and cotangents are [batch, k, ncoef].  k = 0 restores the whole sky.  At
 * most 64 masks here; an empty mask or non-finite weight fails.  Drops the
 * forward state. */
int stlp_set_masks(StlHandle* h, const double* weights, size_t k);
/* Spin-2 front end: not implemented here (lib.rs:625-643). */
int stlp_spin_fields(StlHandle* h, const double* x, size_t n, size_t level, double* qu, double* chi);

/* Symbols of the reference's other handles (patches, spectra, U-Net,
 * training; lib.rs:745-1492), bound eagerly by native.py: they fail with
 * "not implemented" (NULL / -1). */
void* stlp_patch_new(const char* config);
void stlp_patch_free(void* h);
char* stlp_patch_info(void* h);
int stlp_pixel_vectors(size_t nside, const int64_t* ids, size_t n, double* out);
int stlp_patch_select(void* h, const double* x, size_t n, double a, double b, size_t* out);
int stlp_patch_ids(void* h, size_t k, int64_t* out);
int stlp_patch_core_pixels(void* h, size_t* out);
int stlp_patch_forward(void* h, const double* x, size_t n, uint64_t g, bool moments, double* out, double* var);
int stlp_patch_vjp(void* h, const double* cot, size_t n, uint64_t g, double* out);
int stlp_patch_forward_device(void* h, uint64_t x, size_t n, uint64_t g, bool moments, double* out, double* var);
int stlp_patch_vjp_device(void* h, const double* cot, size_t n, uint64_t g, uint64_t out);
void* stlp_spec_new(const char* config);
void stlp_spec_free(void* h);
char* stlp_spec_info(void* h);
int stlp_spec_set_masks(void* h, const double* w, size_t k);
int stlp_spec_forward(void* h, const double* x, size_t n, uint64_t g, uint64_t m, uint64_t o, double* out);
int stlp_spec_vjp(void* h, const double* cot, size_t n, uint64_t g, uint64_t m, uint64_t o, double* out);
void* stlp_unet_new(const char* config);
void stlp_unet_free(void* h);
char* stlp_unet_info(void* h);
int stlp_unet_get_params(void* h, float* out);
int stlp_unet_set_params(void* h, const float* p, size_t n);
int stlp_unet_get_grads(void* h, float* out);
int stlp_unet_zero_grad(void* h);
int stlp_unet_adam(void* h, double lr);
int stlp_unet_set_cond(void* h, const float* c, size_t n);
int stlp_unet_apply(void* h, const double* x, size_t n, double* a, double* b, double* c);
int stlp_train_set_masks(void* h, const double* w, size_t k);
int stlp_train_set_common(void* h, const double* v);
int stlp_train_set_field(void* h, uint32_t f, const double* a, const double* b, const double* c, const double* d,
                         const double* e, const double* g, size_t n, double w);
int stlp_train_bias(void* h, uint32_t f, const uint64_t* a, const double* b, size_t n);
int stlp_train_loss(void* h, uint32_t f, bool b, double* out);
int stlp_train_step(void* h, uint32_t f, double lr, double* out);
int stlp_train_live(void* h, uint32_t f, const uint64_t* a, const double* b, size_t n, const uint64_t* c,
                    const double* d, size_t m, bool e, double w, double* out);
/* Returns a JSON report string in the reference (crates/stl-bridge/src/unet.rs:1477). */
char* stlp_train_lbfgs(void* h, const uint32_t* f, size_t n, const char* options);
int stlp_train_set_joint(void* h, const double* a, const double* b, uint32_t c, uint32_t d, const char* e, double w);
int stlp_train_front(void* h, const double* x, size_t n, bool b, double* out);

/* All-GPU synthesis (no counterpart symbol in the reference; it is the
 * reference's stl-synth objective and L-BFGS, crates/stl-mpi/src/synthesis.rs
 * :207-268 with crates/stl-core/src/lbfgs.rs:356-465, run on device vectors):
 * L-BFGS of loss = sum_k w_k |S_k(x) - target_k|^2 (cotangent 2 w (S - target))
 * from the device maps x_dev (float64, n = batch * 12 nside^2, updated in
 * place), target interleaved (re, im) [batch, K, ncoef] (K = max(masks, 1)),
 * weights [K, ncoef] (one per mask and coefficient, the layout of
 * stlp_native_pixels; the same for every map) or NULL (all 1), options as
 * stlp_lbfgs.  Option "ensemble": true makes the batch one ensemble (the
 * reference's stl-synth --ensemble, synthesis.rs:362-470 without its scatter
 * term): loss = sum_k w_k |mean over the members of S_k - target_k|^2 with
 * target one row [K, ncoef].  Option "history": "evaluations" (default: the
 * loss of every evaluation goes to loss_history, up to cap) or "iterations"
 * (loss_history[0] the start's loss, then the loss at each iteration's
 * accepted step; NaN after the last).  Returns the number of evaluations or
 * -1. */
int64_t stlx_synthesize(StlHandle* h, uint64_t x_dev, size_t n, const double* target, const double* weights,
                        const char* options, double* loss_history, size_t cap);

/* Ensembles of maps of uneven nside (no counterpart in the reference; its
 * ensemble driver loops over NativeEngine handles): groups g of batch[g] maps
 * at nside[g] with j[g] scales, evaluated together, each forward and each VJP
 * one HIP graph per group on concurrent streams (operators shared between the
 * groups' levels of equal nside).  options: precision ("mixed" | "fp64"), device,
 * batch_nside (as stlp_new's, for every group). */
typedef struct StlEnsemble StlEnsemble;
StlEnsemble* stlx_ensemble_new(const int32_t* nside, const int32_t* j, const int32_t* batch, size_t groups,
                               const char* options);
void stlx_ensemble_free(StlEnsemble* e);
/* Coefficients of each group's plan (groups size_t). */
int stlx_ensemble_ncoef(StlEnsemble* e, size_t* out);
/* x_dev[g]: float64 device maps [batch[g], 12 nside[g]^2]; out: interleaved
 * (re, im) statistics concatenated over groups [g][batch][ncoef_g]; var the
 * same without the pairs when moments.  0 or -1. */
int stlx_ensemble_forward_device(StlEnsemble* e, const uint64_t* x_dev, size_t groups, bool moments, double* out,
                                 double* var);
/* Map cotangents of Re sum(conj(cot) values) after a forward: cot as out
 * (n complex values), gradients to the device arrays grad_dev[g]. */
int stlx_ensemble_vjp_device(StlEnsemble* e, const double* cot, size_t n, const uint64_t* grad_dev, size_t groups);

/* Test hooks of this library (no counterpart in the reference): one level's
 * operators of a handle applied to device arrays, for operator-level parity
 * tests.  kind: 0 conv forward (x [rows][N] -> y [rows][4][N] complex),
 * 1 conv adjoint (gy -> gx), 2 decimation forward level -> level + 1,
 * 3 decimation adjoint.  Arrays are float64 on the device. */
int stlx_apply_operator(StlHandle* h, int kind, int level, int rows, uint64_t in_dev, uint64_t out_dev);
/* stlx_apply_operator without any allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * The staging of the operator's input and output lives in a static device
 * scratch of 64 MiB per GPU (held from the library's load until the process
 * ends); an application whose staging exceeds it fails with an error naming
 * the sizes.  Every call on one GPU shares that scratch: calls that overlap
 * on one GPU race.  Results equal stlx_apply_operator's bit for bit. */
int stlx_apply_operator_noalloc(StlHandle* h, int kind, int level, int rows, uint64_t in_dev, uint64_t out_dev);

/* Measurement hook (no counterpart in the reference): mean milliseconds of one
 * copy of `bytes` (a multiple of 16; 16-byte aligned device addresses) from
 * src_dev to dst_dev by this library's hand-written stream kernel, over `reps`
 * copies after a warm-up, to *ms.  The attainable HBM rate is
 * 2 * bytes / ms.  0 or -1. */
int stlx_stream_copy(uint64_t src_dev, uint64_t dst_dev, size_t bytes, int reps, double* ms);

/* Map files and initial maps (no counterpart in the reference's C ABI: they
 * restate the Rust I/O of its command-line tools, crates/stl-mpi/src/io.rs:18-520,
 * and the initial-map draws of crates/stl-cli/src/lib.rs:523-540).  Maps are
 * full-sky NESTED float64; the format follows the extension: ".fits" / ".fit"
 * HEALPix binary tables (float32 "E" or float64 "D" column; RING tables are
 * reordered to NESTED on reading), anything else ".npy" (read: f8, f4, i8, i4
 * of either byte order; written: little-endian f8).  0 or -1. */
/* Pixels of a map file, read from its header only (io.rs:370-379). */
int stlx_map_len(const char* path, size_t* out);
/* The whole map; n must be its length (io.rs:227-233). */
int stlx_read_map(const char* path, double* out, size_t n);
/* NESTED pixels [start, end), read at their offset (io.rs:387-416). */
int stlx_read_map_range(const char* path, size_t start, size_t end, double* out);
/* A .npy array flattened (C order) with its shape (io.rs:27-49).  out NULL:
 * size query, the rank to *ndim and, when shape is not NULL, up to *ndim
 * dimensions to shape.  Otherwise n must be the element count. */
int stlx_read_npy(const char* path, double* out, size_t n, size_t* shape, size_t* ndim);
/* Little-endian float64 .npy of the given shape (io.rs:52-78). */
int stlx_write_npy(const char* path, const double* values, const size_t* shape, size_t ndim);
/* A map file: HEALPix FITS (column name, dbl: "D" else "E") or a flat .npy
 * (io.rs:99-170, 52-55). */
int stlx_write_map(const char* path, const double* values, size_t n, const char* column, bool dbl);
/* A map written by slices (io.rs:431-520): headers and a file sized for n
 * pixels, then pixel ranges [start, start + n) at their offsets.  dbl must
 * be the same in both calls (FITS "D" or "E"; .npy is always f8). */
int stlx_sliced_create(const char* path, size_t n, const char* column, bool dbl);
int stlx_sliced_write(const char* path, bool dbl, size_t start, const double* values, size_t n);
/* Standard normal draws of pixels first .. first + n - 1 for seed:
 * splitmix64 of the seed and the pixel id, Box-Muller of two 53-bit uniforms
 * (lib.rs:523-540); the same map whatever the decomposition. */
int stlx_pixel_normal(uint64_t seed, int64_t first, size_t n, double* out);

#ifdef __cplusplus
}
#endif

#endif /* STL_BRIDGE_H */
