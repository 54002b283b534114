/* simd_bigint_amd.h -- C ABI of the MI355X (gfx950) large-operand multiply.
 *
 * Drop-in for the reference's mpn-style multiply path: the entries below
 * carry the reference's names, argument meaning, length conventions and
 * return codes. Operands are little-endian u64 limb arrays passed as
 * pointer + length. All arithmetic on the GPU is exact integer work (a
 * five-prime NTT over primes below 2^31, Garner CRT and a carry-lookahead
 * scan); results are bit-identical to the reference's.
 *
 * Host-pointer entries are synchronous: they copy the operands to the
 * device, multiply there, copy the product back and return. Device-pointer
 * entries (sba_gpu_*) take device buffers and a HIP stream and only
 * enqueue work.
 *
 * Errors. The int entries of the reference return 1 or 0 (0: outside the
 * band, include/ntmul.h:19). A HIP failure inside a host entry (device or
 * pinned allocation, copy, launch) is reported on stderr and never aborts:
 * nt_* / p30_* / p50_* then return 0 with rp unspecified; the void entries
 * simd_mpn_mul / divrem_u64* finish that product on the library's CPU tier
 * and count it (sba_gpu_fallbacks). A process without any HIP device
 * aborts in a host entry that needs the GPU (the library never runs the
 * GPU path on the CPU unannounced), and host memory exhaustion aborts, as
 * the reference's arena does (include/scratch.h:66-70). Device entries
 * return the codes below.
 *
 * Device memory. The first GPU call on a device builds the root tables of
 * the two prime sets with primes below 2^30: 2 x 5 primes x (2^22 + 2^21)
 * x 4 B = 240 MiB; the wide set's, 2 x 5 x 2^24 x 4 B = 640 MiB, are built
 * by the first call that can use them (a transform above 2^23 points, or
 * any call after sba_set_prime_set_h(0)). All are held until the process
 * ends (one table per prime serves every transform length of its set). Entries without a
 * caller workspace add a grow-only workspace of about 40 B per transform
 * point of the largest product seen (2^23 points: 336 MiB).
 *
 * Host resources. The host-memory entries stage their copies through
 * grow-only pinned buffers (the operands and the product of the largest
 * call seen), and the first staged copy of at least 1 MiB starts 7 copy
 * threads that stay, idle, until the process ends. A forked child copies
 * on its calling thread.
 *
 * Citations are files of the reference repository (iamstupid/simd_bigint).
 */
#ifndef SIMD_BIGINT_AMD_H
#define SIMD_BIGINT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes of the int-valued entries */
#define SBA_OK 1             /* product written */
#define SBA_BAND 0           /* outside the engine's band (reference: return 0) */
#define SBA_ERR_HIP (-1)     /* a HIP runtime call failed (message on stderr) */
#define SBA_ERR_WORKSPACE (-2) /* caller workspace missing or too small */
#define SBA_ERR_NODEV (-3)   /* no HIP device: the library never computes on the CPU */
#define SBA_ERR_TOO_LARGE (-4) /* beyond the allocation-free entry's static scratch */

/* The reference's scratch arena type (include/scratch.h:248-256). The _r
 * entries accept the caller's arena for signature compatibility; device
 * workspace comes from the library's per-device grow-only pool, the same
 * no-per-call-allocation rule the arena enforces. May be NULL. */
struct scratch;

/* ---------------------------------------------------------------------- */
/* NTT dispatcher: replaces include/ntmul.h:47-79.                         */
/* rp[0 .. an+bn) = {ap,an} * {bp,bn}. Returns 1, or 0 exactly outside the */
/* reference's band (sba_ref_nt_band: its p50 band, include/p50.h:1020-    */
/* 1046) and on a HIP failure. rp must not overlap the inputs.             */
/* ---------------------------------------------------------------------- */
int nt_mul(uint64_t *rp, const uint64_t *ap, ptrdiff_t an, const uint64_t *bp,
           ptrdiff_t bn);                                   /* ntmul.h:71-74 */
int nt_sqr(uint64_t *rp, const uint64_t *ap, ptrdiff_t an); /* ntmul.h:76-78 */
int nt_mul_r(uint64_t *rp, const uint64_t *ap, ptrdiff_t an,
             const uint64_t *bp, ptrdiff_t bn,
             struct scratch *sc);                           /* ntmul.h:47-58 */
int nt_sqr_r(uint64_t *rp, const uint64_t *ap, ptrdiff_t an,
             struct scratch *sc);                           /* ntmul.h:60-69 */

/* Engine entries with the reference engines' bands (include/p30.h:1847,
 * 1940,1950,2018 and include/p50.h:1020,1096,1106,1176). Both run the same
 * GPU engine; p30_* additionally return 0 outside p30's band
 * (min(an,bn) > 2,344,425 or ceil((an+bn-1)/16) > 2^18, p30.h:1851-1858). */
int p30_mul(uint64_t *rp, const uint64_t *ap, ptrdiff_t an, const uint64_t *bp,
            ptrdiff_t bn);
int p30_sqr(uint64_t *rp, const uint64_t *ap, ptrdiff_t an);
int p30_mul_r(uint64_t *rp, const uint64_t *ap, ptrdiff_t an,
              const uint64_t *bp, ptrdiff_t bn, struct scratch *sc);
int p30_sqr_r(uint64_t *rp, const uint64_t *ap, ptrdiff_t an, struct scratch *sc);
int p50_mul(uint64_t *rp, const uint64_t *ap, ptrdiff_t an, const uint64_t *bp,
            ptrdiff_t bn);
int p50_sqr(uint64_t *rp, const uint64_t *ap, ptrdiff_t an);
int p50_mul_r(uint64_t *rp, const uint64_t *ap, ptrdiff_t an,
              const uint64_t *bp, ptrdiff_t bn, struct scratch *sc);
int p50_sqr_r(uint64_t *rp, const uint64_t *ap, ptrdiff_t an, struct scratch *sc);

/* ---------------------------------------------------------------------- */
/* Public multiply with threshold dispatch: replaces include/u64cvt.h:94-  */
/* 132. rp[0 .. an+bn) = product; operands in either order; never writes   */
/* past an+bn; zero operands give zero. Sizes below the GPU crossover run  */
/* the library's own CPU basecase / Karatsuba (in the roles of the         */
/* reference's ladder, include/mul.h:764-795; no Toom level), with a       */
/* per-thread grow-only scratch; above it, the GPU NTT.                    */
/* ---------------------------------------------------------------------- */
void simd_mpn_mul(uint64_t *rp, const uint64_t *ap, uint64_t an,
                  const uint64_t *bp, uint64_t bn);

/* ---------------------------------------------------------------------- */
/* Division: replaces include/divrem.h:218-221 and include/dc_div.h:154-   */
/* 157. Q = floor(N/D) into qp (>= nn64-dn64+1 limbs), R = N mod D into rp */
/* (>= dn64 limbs), trimmed lengths (>= 1) into *qn64_out / *rn64_out.     */
/* dp[dn64-1] != 0. Newton reciprocal on the GPU multiply.                 */
/* ---------------------------------------------------------------------- */
void divrem_u64(uint64_t *qp, uint64_t *rp, const uint64_t *np, uint64_t nn64,
                const uint64_t *dp, uint64_t dn64, uint64_t *qn64_out,
                uint64_t *rn64_out);
void divrem_u64_dc(uint64_t *qp, uint64_t *rp, const uint64_t *np,
                   uint64_t nn64, const uint64_t *dp, uint64_t dn64,
                   uint64_t *qn64_out, uint64_t *rn64_out);

/* ---------------------------------------------------------------------- */
/* Device-resident entries (no counterpart in the reference, which has no  */
/* device): device pointers, a HIP stream (hipStream_t passed as void*,    */
/* NULL = default stream). Enqueue only. Return SBA_OK / SBA_BAND / < 0.   */
/* Without a caller workspace, calls share the device's grow-only          */
/* workspace: two of them on different streams must not overlap in time   */
/* (the _ws variants, each with its own workspace, may).                   */
/* ---------------------------------------------------------------------- */
int sba_gpu_mul_dev(uint64_t *rp, const uint64_t *ap, int64_t an,
                    const uint64_t *bp, int64_t bn, void *stream);
int sba_gpu_sqr_dev(uint64_t *rp, const uint64_t *ap, int64_t an, void *stream);
/* same with a caller-provided device workspace of ws_bytes >=
 * sba_gpu_workspace_bytes(an, bn) bytes */
int sba_gpu_mul_dev_ws(uint64_t *rp, const uint64_t *ap, int64_t an,
                       const uint64_t *bp, int64_t bn, void *ws, size_t ws_bytes,
                       void *stream);
int sba_gpu_sqr_dev_ws(uint64_t *rp, const uint64_t *ap, int64_t an, void *ws,
                       size_t ws_bytes, void *stream);
size_t sba_gpu_workspace_bytes(int64_t an, int64_t bn);
/* Allocation-free multiply. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same contract as
 * sba_gpu_mul_dev (bp == ap with bn == an runs the square), but the call
 * makes no hipMalloc / hipFree / allocating library call: its spectra,
 * carry maps and root tables live in a static device array of
 * SBA_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends. Products
 * of up to 2^SBA_NOALLOC_MAX_LG coefficients (an + bn - 1), always one
 * transform; larger ones return SBA_ERR_TOO_LARGE. The first call on a
 * device builds that device's tables and waits for them. All calls on one
 * GPU share its array: calls that overlap in time on one GPU (different
 * streams, threads) race and must be serialised by the caller. */
#define SBA_NOALLOC_MAX_LG 20
#define SBA_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int sba_gpu_mul_dev_noalloc(uint64_t *rp, const uint64_t *ap, int64_t an,
                            const uint64_t *bp, int64_t bn, void *stream);
/* 1 when the device entries accept (an, bn): min(an, bn) within the CRT
 * cap and either an + bn - 1 <= 2^25 (one transform) or a chunked product
 * (the longer operand cut into chunks against the shorter one's spectrum,
 * an + bn <= 2^29). Batches take only the one-transform band. */
int sba_gpu_in_band(int64_t an, int64_t bn);
/* how a single product of (an, bn) runs: *nchunks == 0 -> one transform;
 * else chunks of *chunk limbs of the longer operand against the shorter
 * one's spectrum of length 2^*Lc */
int sba_gpu_chunking(int64_t an, int64_t bn, int sqr, int *Lc, uint64_t *chunk,
                     uint64_t *nchunks);
/* L when a single product of (an, bn) runs at the radix-3 length 3 * 2^L
 * (2^(L+1) < an+bn-1 <= 3 * 2^L, sba_get_r3_min_lg() <= L <= 22, not
 * chunked), else 0. Batch jobs take radix-3 classes from L = 13. */
int sba_gpu_radix3(int64_t an, int64_t bn, int sqr);
/* the schoolbook route of a product: 1 the one-workgroup schoolbook (an +
 * bn <= 2048 limbs, see sba_set_school_max), 2 the short x long schoolbook
 * (the shorter operand <= 512 limbs and an * bn <= ratio / 16 * N log2 N,
 * N the NTT length the product would get; sba_set_school_long_ratio, 0 turns
 * it off, default measured on MI355X), 0 the NTT */
int sba_gpu_school(int64_t an, int64_t bn, int sqr);
void sba_set_school_long_ratio(uint32_t sixteenths);
uint32_t sba_get_school_long_ratio(void);
/* smallest radix-3 row length 2^L for single products (default 16, measured
 * on MI355X; batches always from 13) */
void sba_set_r3_min_lg(int L);
int sba_get_r3_min_lg(void);
/* the prime sets of power-of-two transforms up to 2^23 and of radix-3
 * transforms up to 3 x 2^22: 1 (default) three primes below 2^30, whose
 * butterflies take 8 instructions instead of 10, and two below 2^31 (one
 * set for each kind of length); 0 the five primes below 2^31 that every
 * other transform uses. Results are the same either way; batches keep the
 * choice in force when they were created. */
void sba_set_prime_set_h(int on);
int sba_get_prime_set_h(void);
/* 1 when a single product of (an, bn) runs truncated (a power-of-two length
 * 2^*L whose merged last pass transforms only the first *rows tiles of
 * 2^*last residues, rows = ceil((an+bn-1) / 2^last), and whose inverse
 * column pass is the truncated one), else 0 */
int sba_gpu_truncated(int64_t an, int64_t bn, int sqr, int *L, int *last, int *rows);
/* The carry stage between blocks of sba_gpu_carry_block_limbs() limbs on
 * its own (no counterpart in the reference; it lets a test drive the stage
 * with chosen carry chains): rp[0..Z) holds limbs each finished with block
 * carry-in 0, bmaps[b] the 2-state carry map of block b (bit c = carry out
 * for carry in c); adds the carries between blocks in place. Enqueue only;
 * SBA_OK, SBA_BAND (a null pointer, Z = 0 or Z > 2^29) or < 0. */
uint32_t sba_gpu_carry_block_limbs(void);
int sba_gpu_carry_blocks(uint64_t *rp, const uint32_t *bmaps, uint64_t Z, void *stream);
uint64_t sba_gpu_min_operand_cap(void);
int sba_gpu_max_lg(void);
int sba_gpu_device_count(void);

/* ---------------------------------------------------------------------- */
/* Batches (no counterpart in the reference, which has no batch entry):    */
/* many independent multiplies of uneven sizes on device buffers. A job    */
/* with bp == NULL is a square of ap. sba_batch_create groups the jobs in  */
/* size classes, allocates their workspace once and (use_graph != 0)       */
/* captures the whole launch sequence into a hipGraph; sba_batch_run       */
/* replays it on a stream (enqueue only). Jobs keep their pointers: rerun  */
/* after rewriting the operand buffers. *status: SBA_OK, SBA_BAND (a job   */
/* outside the band) or < 0.                                               */
/* ---------------------------------------------------------------------- */
typedef struct {
  uint64_t *rp;            /* an+bn limbs (2an for a square) */
  const uint64_t *ap;
  int64_t an;
  const uint64_t *bp;      /* NULL: square of ap */
  int64_t bn;
} sba_job;
typedef struct sba_batch sba_batch;
sba_batch *sba_batch_create(const sba_job *jobs, int njobs, int use_graph, int *status);
int sba_batch_run(sba_batch *b, void *stream);
int sba_batch_classes(const sba_batch *b);
void sba_batch_destroy(sba_batch *b);
/* one-shot launch of a batch into a caller workspace of >=
 * sba_batch_ws_bytes(jobs, n) bytes: no allocation, no graph (the serving
 * loop's entry). Returns after enqueueing; the workspace is in use until the
 * stream passes this point. */
size_t sba_batch_ws_bytes(const sba_job *jobs, int njobs);
int sba_batch_launch_ws(const sba_job *jobs, int njobs, void *ws, size_t ws_bytes,
                        void *stream);
/* the same jobs through the plain per-call entries, one after another */
int sba_gpu_mul_loop(const sba_job *jobs, int njobs, void *stream);

/* ---------------------------------------------------------------------- */
/* Plans and autotuning (no counterpart in the reference: its transform    */
/* shape is fixed per size by code, include/p30.h:1847-1938). A plan cuts  */
/* a transform of length 2^L into passes:                                  */
/*   last   levels of the last pass, on contiguous tiles of 2^last         */
/*          residues; last == L (L <= 13) is a single-tile transform,      */
/*          else a first pass over column tiles of 2^13 residues takes the */
/*          other L - last levels (last in 11..13, 1 <= L-last <= 12)      */
/*   bm     widest register chunk in a tile: radix-2^bm stages (3..5)      */
/*   split  first forward pass one prime per block (1) or all five        */
/*          primes per block (0)                                           */
/* The library keeps one plan per (L, square?, bucket of the class's job   */
/* count); every launch (single call, batch class, serving batch) uses the */
/* plan stored for its class at that time. The table starts as (and        */
/* sba_plan_reset restores) the shipped one: the fastest plan per key      */
/* measured on MI355X, the hand-written rule (sba_plan_heuristic: last =   */
/* min(L, 13), bm 5, split 1) where a key was not measured. sba_plan_set   */
/* returns SBA_BAND for an invalid plan.                                   */
/* kind = (square ? 1 : 0) | (radix-3 ? 2 : 0). A radix-3 class has length */
/* 3 * 2^L (L = 13..22; its plans: last 11..13, 1..9 first-pass levels);    */
/* its shipped plans are the power-of-two ones moved into those limits.    */
/* ---------------------------------------------------------------------- */
#define SBA_PLAN_BUCKETS 6
int sba_plan_bucket(int njobs);    /* 1-3 jobs: 0, 4-15: 1, ... >= 1024: 5 */
int sba_plan_set(int L, int kind, int bucket, int last, int bm, int split);
int sba_plan_get(int L, int kind, int bucket, int *last, int *bm, int *split);
int sba_plan_heuristic(int L, int *last, int *bm, int *split);
void sba_plan_reset(void);
/* valid plans of class (L, kind) as (last, bm, split) triples; returns the
 * count (writes at most max triples) */
int sba_plan_candidates(int L, int kind, int *plans, int max);
/* time one size class (all jobs the same L and kind) under a plan: one
 * warm-up, then reps launches between HIP events on the stream; *ms = mean
 * per launch. Workspace as sba_batch_launch_ws. Synchronizes the stream. */
int sba_class_time(const sba_job *jobs, int njobs, int last, int bm, int split, int reps,
                   void *ws, size_t ws_bytes, void *stream, float *ms);
/* the size class of a batch job: its kind (as above) and L (transform
 * length 2^L, or 3 * 2^L for radix-3); SBA_BAND outside the batch band */
int sba_job_class(const sba_job *j, int *L, int *kind);

/* Autotuner: for every size class of the jobs, time candidate plans on the
 * class's own jobs (reps launches each, products written to the jobs' rp)
 * and store the fastest for (L, kind, bucket of the class's job count).
 *   SBA_TUNE_DEFAULT     no change
 *   SBA_TUNE_EXHAUSTIVE  every valid plan is timed
 *   SBA_TUNE_MODEL       the loaded cost model ranks the plans; the topk
 *                        best-ranked are timed (topk = 1: none is, the
 *                        model's first choice is stored). The model covers
 *                        power-of-two classes; radix-3 classes keep their
 *                        plan in this mode.
 * Workspace as sba_batch_launch_ws. *measured = plan timings made.
 * SBA_ERR_WORKSPACE: workspace too small, or no model loaded for MODEL. */
#define SBA_TUNE_DEFAULT 0
#define SBA_TUNE_EXHAUSTIVE 1
#define SBA_TUNE_MODEL 2
int sba_autotune(const sba_job *jobs, int njobs, int mode, int topk, int reps, void *ws,
                 size_t ws_bytes, void *stream, int *measured);

/* Cost model: an MLP (SBA_CM_IN -> 32 -> 32 -> 1, ReLU) over plan
 * features, predicting a plan's standardised log launch time relative to
 * the other plans of its class (the autotuner only ranks plans).
 * Parameters: W1[32][16], b1[32], W2[32][32], b2[32], w3[32], b3
 * (SBA_CM_NPARAM floats). The training entries take device pointers and
 * enqueue on the stream:
 *   sba_cm_forward   y[n] = f(x[n][SBA_CM_IN])
 *   sba_cm_backward  grad = d/dparams of (1/n) sum (f(x) - t)^2; ws holds
 *                    sba_cm_ws_floats(n) floats, predictions in ws[0..n)
 *   sba_cm_adam      one Adam step (b1 0.9, b2 0.999, eps 1e-8), step >= 1
 * sba_cm_features writes the raw feature row of (class, plan) (host);
 * sba_cm_load installs host parameters plus the feature / target
 * standardisation for the autotuner; sba_cm_predict returns its estimate
 * as a relative cost (comparable between plans of one class). */
#define SBA_CM_IN 16
#define SBA_CM_NPARAM 1633
int sba_cm_nparams(void);
size_t sba_cm_ws_floats(int n);
int sba_cm_forward(const float *params, const float *x, int n, float *y, void *stream);
int sba_cm_backward(const float *params, const float *x, const float *t, int n, float *grad,
                    float *ws, void *stream);
int sba_cm_adam(float *params, const float *grad, float *m, float *v, int step, float lr,
                void *stream);
int sba_cm_features(int L, int sqr, int njobs, double fill, int last, int bm, int split,
                    float *f);
int sba_cm_load(const float *params, const float *fmean, const float *fstd, float ymean,
                float ystd);
int sba_cm_loaded(void);
int sba_cm_predict(int L, int sqr, int njobs, double fill, int last, int bm, int split,
                   void *stream, double *ms);

/* ---------------------------------------------------------------------- */
/* Tuning and test hooks (no counterpart in the reference, whose tier      */
/* thresholds are compile-time macros, include/u52-mparam.h:1-19).         */
/* ---------------------------------------------------------------------- */
/* simd_mpn_mul and the division send products with an+bn >= this many
 * limbs to the GPU, smaller ones to the CPU Karatsuba. */
void sba_set_gpu_threshold(uint64_t min_total_limbs);
uint64_t sba_get_gpu_threshold(void);
/* products with an + bn <= 2048 limbs run on the schoolbook kernel (one
 * workgroup per product) instead of the NTT when an * bn <= max_product *
 * N / 1024 in a batch and half that in a single call (N = the NTT length
 * the product would get, at least 1024); 0 turns that tier off. The
 * default, 32768, is the MI355X-measured crossover. */
void sba_set_school_max(uint64_t max_product);
uint64_t sba_get_school_max(void);
/* truncation: a two-pass product (or batch class) of length >= 2^min_lg
 * whose shortest product needs at most max_fill_pct percent of its
 * last-pass tiles runs truncated; products of such lengths that fill at
 * most r3_min_pct percent of their power-of-two length take that truncated
 * length instead of the radix-3 length. max_fill_pct 0 turns truncation off.
 * Defaults, measured on MI355X: 85, 0, 21. */
void sba_set_trunc(int max_fill_pct, int r3_min_pct, int min_lg);
void sba_get_trunc(int *max_fill_pct, int *r3_min_pct, int *min_lg);
/* crossover tuner: times balanced products of each total in the ladder
 * that sba_tune_sizes lists (host buffers; the CPU multiply against the GPU
 * path with its copies, best of 5) and returns the smallest total from
 * which the GPU wins at every larger ladder size (twice the largest if it
 * never does); apply != 0 also sets it as the threshold. cpu_ms / gpu_ms
 * (may be NULL) receive the times, one per ladder size. */
int sba_tune_sizes(uint64_t *totals, int max);
uint64_t sba_tune_gpu_threshold(int apply, double *cpu_ms, double *gpu_ms);
/* per-kernel timing of the GPU path (HIP events on the launch stream):
 * enable, run, synchronize, collect; then read per-kernel totals (name,
 * milliseconds, algorithmic bytes, launches) by index until it returns 0. */
void sba_gpu_profile_enable(int on);
int sba_gpu_profile_collect(void);
void sba_gpu_profile_reset(void);
int sba_gpu_profile_get(int idx, const char **name, double *ms, double *bytes,
                        uint64_t *count);
/* radix-2 butterflies per second that the current device issues from
 * registers alone (no memory traffic): the VALU ceiling the transform
 * kernels are measured against. SBA_OK, or an error code. */
int sba_gpu_probe_butterflies(double *per_s);
/* 1 when the reference's nt_mul would accept (an, bn) (its p50 band,
 * include/p50.h:1020-1046); nt_mul / nt_sqr / p50_* return 0 exactly
 * outside it (or on a HIP failure). */
int sba_ref_nt_band(int64_t an, int64_t bn);
/* the CPU tier alone (basecase + Karatsuba), either operand order */
void sba_cpu_mul(uint64_t *rp, const uint64_t *ap, uint64_t an,
                 const uint64_t *bp, uint64_t bn);
/* products that simd_mpn_mul or the division finished on the CPU tier
 * after a HIP failure, since the process started */
uint64_t sba_gpu_fallbacks(void);
/* failure injection for tests of the error contract: the next `count`
 * GPU multiplies of the host entries fail before touching the device
 * (SBA_FAIL_ENTRY), or the next `count` device / pinned staging
 * allocations of the host entries fail (SBA_FAIL_ALLOC); count 0 disarms */
#define SBA_FAIL_ENTRY 0
#define SBA_FAIL_ALLOC 1
void sba_test_inject_failure(int where, int count);

#ifdef __cplusplus
}
#endif
#endif
