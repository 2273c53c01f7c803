/*
 * bytecrt_probes.h -- measurement probes beside the ByteCRT product library
 * (fherma-bytecrt_amd/lib/libbytecrt_probes.so).  None has a counterpart in
 * the reference: they measure the two ceilings the product's kernels are held
 * against on gfx950, with code of this repository rather than a runtime blit.
 *
 * Status codes are those of bytecrt.h (BCRT_OK, BCRT_EINVAL, BCRT_ERUNTIME).
 */
#ifndef BYTECRT_PROBES_H
#define BYTECRT_PROBES_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* HBM streaming on `bytes` (a multiple of 16) of device memory, enqueued on
 * `stream` (hipStream_t; NULL = default), with d_src and d_dst both `bytes`
 * long.  mode 0: copy d_src -> d_dst (read and write); 1: read d_src only
 * (each thread's xor of what it read lands in the first 16 * threads bytes of
 * d_dst, so bytes must cover one 16-byte word per thread); 2: write d_dst
 * only (a fill).  16-byte accesses, four in flight per thread (non-temporal
 * for copy and read), a grid of 32 workgroups of 256 threads per CU. */
int bcrtp_hbm_stream(int mode, const void* d_src, void* d_dst, size_t bytes, void* stream);

/* Butterfly throughput of one exact modular-multiplication form: each of
 * `threads` threads (a multiple of 256) holds 16 residues mod p and runs
 * `rounds` radix-16 rounds (4 levels, 32 Cooley-Tukey butterflies x + w y,
 * x - w y each) in registers with the 15 twiddles tw15 (values in [0, p)),
 * then writes its 16 canonical residues to d_out[16 * thread + k].
 * form 0: FP64 (TwoProduct, rounded quotient, exact remainder; the product
 *         kernels' arithmetic; values balanced again after every round);
 * form 1: 32-bit Shoup (precomputed floor(w 2^32 / p), v_mul_hi_u32);
 * form 2: 32-bit Montgomery (R = 2^32);
 * form 3: FP64 in five operations for p - 1 = k 2^s (q = rint(x w/p),
 *         fma(x, w, -(q p - q)) - q; needs 3 p k < 2^53, else BCRT_EINVAL);
 * form 4: form 3 with each pair of butterflies written step by step,
 *         interleaved;
 * form 5: form 0 with the product x w formed by an FMA with a -0.0 addend
 *         (v_fma_f64 instead of v_mul_f64; the same doubles).
 * 2^30 < p < 2^31 prime.  Every form computes the same residues.
 * lds_bytes (<= 65536) of dynamic LDS per 256-thread workgroup, which only
 * limits how many waves share a SIMD (0: registers decide). */
int bcrtp_butterflies(int form, uint32_t* d_out, size_t threads, uint32_t p, const uint32_t* tw15, uint32_t rounds,
                      uint32_t lds_bytes,
                      void* stream);

/* Overlap of matrix-core and vector work on one SIMD: a grid of two 512-thread
 * workgroups per CU, whose waves 0..3 (one per SIMD) each run 4 x mfma_iters
 * v_mfma_i32_32x32x32_i8 (four independent chains) and whose waves 4..7 run
 * 8 x valu_iters operations of one kind (eight independent chains): kind 0
 * FP64 fma, 1 FP32 fma, 2 int32 24-bit multiply-add.  mode 1 runs only the
 * MFMA waves, 2 only the VALU waves, 3 both; a running thread writes one word
 * d_out[512 * workgroup + thread] (out_words >= bcrtp_issue_overlap_words()).
 * Timing modes 1, 2 and 3 shows whether the two kinds of work overlap (mode 3
 * near the larger of 1 and 2) or take turns (near their sum). */
int bcrtp_issue_overlap(int mode, int kind, uint32_t mfma_iters, uint32_t valu_iters, int32_t* d_out,
                        size_t out_words, void* stream);
size_t bcrtp_issue_overlap_words(void);

#ifdef __cplusplus
}
#endif

#endif /* BYTECRT_PROBES_H */
