/*
 * mpsh.h -- C ABI of the MI355X block-sparse U(1) contraction engine.
 *
 * The entry points replace the Julia hot path of amiragha/mps:
 *   SymTensors.contract / permutelegs / fuselegs / splitleg / SymMatrix,
 *   TensorNetAlgs._applymps2site / _mpsupdateleft / _mpsupdateright and the
 *   block-wise vector operations KrylovKit's eigsolve calls on SymTensors.
 * Plain pointers and sizes only.  Device pointers are HIP device addresses;
 * `stream` is a hipStream_t (NULL = default stream).  Tensor data is the
 * concatenation of all blocks of a structure in column-major sector order,
 * each block column-major (reference src/symtensor.jl:12-30, src/sector.jl:30-36).
 *
 * Errors: every call returns an mpsh_status; mpsh_last_error() gives a
 * thread-local message.  Plans and structures are immutable once created;
 * one plan may be executed concurrently only on distinct workspaces.
 */
#ifndef MPSH_H
#define MPSH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  MPSH_OK = 0,
  MPSH_ERR_INVALID = 1,         /* malformed arguments (index sets, ranks, NULLs)            */
  MPSH_ERR_SPACE_MISMATCH = 2,  /* "not contractible!"  (src/contract.jl:237-238)              */
  MPSH_ERR_SECTOR_MISMATCH = 3, /* block table inconsistent with the space (src/symtensor.jl:20-26) */
  MPSH_ERR_NOT_SIMILAR = 4,     /* axpy!/dot on non-similar tensors (src/symtensor.jl:291,316) */
  MPSH_ERR_NONFINITE = 5,
  MPSH_ERR_DEVICE = 6,          /* a HIP runtime call failed                                  */
  MPSH_ERR_OOM = 7,
  MPSH_ERR_UNSUPPORTED = 8      /* e.g. rank-1 outputs of contract (src/contract.jl:24-42)   */
} mpsh_status;

typedef enum {
  MPSH_F64 = 0,  /* double storage, FP64 MFMA                                    */
  MPSH_F32 = 1,  /* float storage, FP32 MFMA                                     */
  MPSH_BF16 = 2  /* float storage, BF16 MFMA operands, FP32 accumulation         */
} mpsh_dtype;

const char* mpsh_last_error(void);
const char* mpsh_version(void);

/* ------------------------------------------------------------------------ */
/* Structures: a tensor charge plus U(1) legs, with every allowed sector.    */
/* Replaces SymTensor's sector bookkeeping: VectorSpace (src/vspace.jl:5-18), */
/* _allsectorsandsizes (src/sector.jl:69-90), the SortedDict key table        */
/* (src/sorteddict.jl:1-99).                                                  */
/* ------------------------------------------------------------------------ */
typedef struct mpsh_struct_s* mpsh_struct;

/* legs: isdual[rank]; ncharges[rank]; charges/dims concatenated leg by leg.  */
mpsh_status mpsh_struct_create(int64_t charge, int32_t rank, const uint8_t* isdual, const int32_t* ncharges,
                               const int64_t* charges, const int64_t* dims, mpsh_struct* out);
void mpsh_struct_destroy(mpsh_struct s);
int32_t mpsh_struct_rank(mpsh_struct s);
int64_t mpsh_struct_charge(mpsh_struct s);
int64_t mpsh_struct_nblocks(mpsh_struct s);
int64_t mpsh_struct_numel(mpsh_struct s);
int32_t mpsh_struct_leg_ncharges(mpsh_struct s, int32_t leg); /* leg is 0-based */
mpsh_status mpsh_struct_leg(mpsh_struct s, int32_t leg, uint8_t* isdual, int64_t* charges, int64_t* dims);
/* sectors[nblocks*rank], sizes[nblocks*rank], offsets[nblocks+1] (elements). */
mpsh_status mpsh_struct_blocks(mpsh_struct s, int64_t* sectors, int64_t* sizes, int64_t* offsets);
/* block index of a sector, -1 when absent */
int64_t mpsh_struct_find(mpsh_struct s, const int64_t* sector);

/* ------------------------------------------------------------------------ */
/* Re-legging: permutelegs (src/symreleg.jl:34-51), fuselegs (:79-146),      */
/* splitleg (:161-214), SymMatrix (src/contract.jl:112-171).  Leg numbers and */
/* permutations are 1-based as in the reference.                              */
/* ------------------------------------------------------------------------ */
typedef struct mpsh_releg_s* mpsh_releg;

mpsh_status mpsh_permutelegs_plan(mpsh_struct A, const int32_t* perm, mpsh_releg* plan, mpsh_struct* out);
mpsh_status mpsh_fuselegs_plan(mpsh_struct A, int32_t l, int32_t n, int32_t dual, mpsh_releg* plan,
                               mpsh_struct* out);
mpsh_status mpsh_splitleg_plan(mpsh_struct A, int32_t l, int32_t nspaces, const uint8_t* isdual,
                               const int32_t* ncharges, const int64_t* charges, const int64_t* dims,
                               mpsh_releg* plan, mpsh_struct* out);
mpsh_status mpsh_symmatrix_plan(mpsh_struct A, int32_t nrows, const int32_t* rows, int32_t ncols,
                                const int32_t* cols, mpsh_releg* plan, mpsh_struct* out);
mpsh_status mpsh_releg_execute(mpsh_releg plan, mpsh_dtype dt, const void* src, void* dst, void* stream);
void mpsh_releg_destroy(mpsh_releg plan);

/* ------------------------------------------------------------------------ */
/* contract(A, idxA, B, idxB)  (src/contract.jl:11-47): negative labels are   */
/* contracted (ordered by |label|), positive labels are output slots 1..K.    */
/* Plain path: gather-fuse A and B, one grouped GEMM over all fused charges,  */
/* scatter-split-permute C -- the reference algorithm on the device.          */
/* ------------------------------------------------------------------------ */
typedef struct mpsh_contract_s* mpsh_contract;

mpsh_status mpsh_contract_plan(mpsh_struct A, const int32_t* idxA, mpsh_struct B, const int32_t* idxB,
                               mpsh_contract* plan, mpsh_struct* C);
size_t mpsh_contract_workspace(mpsh_contract plan, mpsh_dtype dt); /* bytes */
int64_t mpsh_contract_flops(mpsh_contract plan);                   /* sum 2*M*N*K of matched blocks */
mpsh_status mpsh_contract_execute(mpsh_contract plan, mpsh_dtype dt, const void* A, const void* B, void* C,
                                  void* workspace, void* stream);
void mpsh_contract_destroy(mpsh_contract plan);

/* ------------------------------------------------------------------------ */
/* Two-site effective Hamiltonian  _applymps2site(v, envL, envR, WL, WR)      */
/* (src/apply.jl:93-105).  Fused path: step-1 grouped GEMM (v^T envL), one    */
/* combined-MPO pass (WL x WR over the middle bond), step-4 grouped GEMM with */
/* K segments over (mpo_r, ket_r).  Output structure and block layout equal   */
/* the reference's four-contract chain.                                       */
/* ------------------------------------------------------------------------ */
typedef struct mpsh_heff_s* mpsh_heff;

mpsh_status mpsh_heff_plan(mpsh_struct v, mpsh_struct envL, mpsh_struct envR, mpsh_struct WL, mpsh_struct WR,
                           mpsh_heff* plan, mpsh_struct* out);
/* stats[0..9]: reference-order FLOPs, reference-order elements moved, step FLOPs x4,
 * step elements x4 (the reference's own four contracts, src/apply.jl:99-103). */
mpsh_status mpsh_heff_stats(mpsh_heff plan, int64_t* stats);
/* stats[0..7] of this engine's own stages (instrumentation, no reference counterpart):
 * T1 elements, T3 elements, combined-MPO elements, MPO-pass FLOPs, step-1 GEMM FLOPs,
 * step-4 GEMM FLOPs, packed-v elements, per-matvec workspace elements. */
mpsh_status mpsh_heff_engine_stats(mpsh_heff plan, int64_t* stats);

/* One-shot application, the drop-in for _applymps2site: workspace of
 * mpsh_heff_workspace() bytes, 16-byte aligned. */
size_t mpsh_heff_workspace(mpsh_heff plan, mpsh_dtype dt);
mpsh_status mpsh_heff_apply(mpsh_heff plan, mpsh_dtype dt, const void* v, const void* envL, const void* envR,
                            const void* WL, const void* WR, void* out, void* workspace, void* stream);

/* Prepared operator for repeated application inside one eigensolve (the
 * environments and MPO of a bond are fixed while Lanczos runs, dmrg.jl:77-82):
 * envL/envR are repacked into 16-byte aligned K-contiguous rows and the
 * combined two-site MPO is formed once; each matvec then packs v and runs
 * the three kernel stages.  Buffers are 16-byte aligned device memory. */
size_t mpsh_heff_prepared_size(mpsh_heff plan, mpsh_dtype dt);
size_t mpsh_heff_apply_workspace(mpsh_heff plan, mpsh_dtype dt);
mpsh_status mpsh_heff_prepare(mpsh_heff plan, mpsh_dtype dt, const void* envL, const void* envR, const void* WL,
                              const void* WR, void* prepared, void* stream);
mpsh_status mpsh_heff_apply_prepared(mpsh_heff plan, mpsh_dtype dt, const void* prepared, const void* v,
                                     void* out, void* workspace, void* stream);
/* The same split by stage for per-kernel timing: stages is a bit mask,
 * 1 = pack v, 2 = step-1 GEMM, 4 = combined-MPO pass, 8 = step-4 GEMM; run in
 * that order; mask 15 equals mpsh_heff_apply_prepared.  Instrumentation,
 * no reference counterpart. */
mpsh_status mpsh_heff_apply_prepared_stages(mpsh_heff plan, mpsh_dtype dt, int32_t stages, const void* prepared,
                                            const void* v, void* out, void* workspace, void* stream);
/* The same application without a workspace and without allocating: no
 * hipMalloc / hipFree or allocating library call, so it may run under HIP graph
 * capture and in loops that must not allocate.  No counterpart in the
 * reference; for callers that cannot allocate.  The per-matvec scratch (packed
 * v, T1, T3, split-K partials) is a fixed device array of
 * MPSH_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; plans whose
 * mpsh_heff_apply_workspace exceeds it are rejected (MPSH_ERR_INVALID), as is a
 * plan whose tables are not yet on the current device (mpsh_heff_prepare for
 * this dtype uploads them).  All calls on one GPU share the array: calls that
 * overlap on one GPU race.  Results are bitwise those of
 * mpsh_heff_apply_prepared. */
#define MPSH_NOALLOC_SCRATCH_BYTES (64u << 20)
mpsh_status mpsh_heff_apply_prepared_noalloc(mpsh_heff plan, mpsh_dtype dt, const void* prepared, const void* v,
                                             void* out, void* stream);
/* Autotuner over the plan's block-size histogram: times every candidate tile
 * variant of the step-1 GEMM and every (tile, split-K) variant of the step-4
 * GEMM on the plan's real problems (reps launches each, HIP events on `stream`,
 * synchronises), keeps the fastest for dtype dt in the plan, and records it
 * under the plan's histogram signature so that later plans with the same
 * signature adopt it without timing (force = 1 re-times).  choice[9] (may be
 * NULL): step-1 BM, BN, KS, 0, step-4 BM, BN, KS, split target, 1 if taken from
 * the signature cache.  Buffers as for mpsh_heff_apply_prepared; the output
 * and workspace contents are garbage afterwards.  No reference counterpart. */
mpsh_status mpsh_heff_tune(mpsh_heff plan, mpsh_dtype dt, const void* prepared, const void* v, void* out,
                           void* workspace, int32_t reps, int32_t force, int32_t* choice, void* stream);
void mpsh_heff_destroy(mpsh_heff plan);

/* ------------------------------------------------------------------------ */
/* Environment updates  _mpsupdateleft(lenv, A, W)  (src/apply.jl:49-58) and    */
/* _mpsupdateright(renv, A, W)  (src/apply.jl:17-25), the reference's three     */
/* contracts each, called from initialenv / dmrg2sitesweep! (src/dmrg.jl:29,33, */
/* :103,141,175).  Fused plan: the H_eff's step-1 grouped GEMM (env x A) and     */
/* combined-MPO pass over one site, then one grouped GEMM against the bra       */
/* dual(A) (A's own data, no conjugation: src/symtensor.jl:201-212).  The right */
/* update runs as the left update of the mirror image A(r, p, l),               */
/* W(w_r, p_out, w_l, p_in).  The plan returns the output structure; outputs    */
/* and block order equal the reference chain's.  FP64 and FP32 (BF16: error).   */
/* stats[4]: reference-chain FLOPs, reference-chain elements moved, this        */
/* engine's GEMM FLOPs, its MPO-pass FLOPs.                                     */
/* ------------------------------------------------------------------------ */
typedef struct mpsh_envupdate_s* mpsh_envupdate;

mpsh_status mpsh_envupdate_left_plan(mpsh_struct lenv, mpsh_struct A, mpsh_struct W, mpsh_envupdate* plan,
                                     mpsh_struct* out);
mpsh_status mpsh_envupdate_right_plan(mpsh_struct renv, mpsh_struct A, mpsh_struct W, mpsh_envupdate* plan,
                                      mpsh_struct* out);
size_t mpsh_envupdate_workspace(mpsh_envupdate plan, mpsh_dtype dt); /* bytes, 16-byte aligned buffer */
mpsh_status mpsh_envupdate_stats(mpsh_envupdate plan, int64_t* stats);
mpsh_status mpsh_envupdate_execute(mpsh_envupdate plan, mpsh_dtype dt, const void* env, const void* A,
                                   const void* W, void* out, void* workspace, void* stream);
void mpsh_envupdate_destroy(mpsh_envupdate plan);

/* ------------------------------------------------------------------------ */
/* Adjoint of E(v, envL, envR) = <v|H_eff v> (H_eff of _applymps2site,          */
/* src/apply.jl:93-105; the reference has no AD): the gradients of the          */
/* un-normalised quadratic form w.r.t. the environments,                        */
/*   gL[a,m,b] = sum v[a,p,s,c] WL[m,q,n,p] WR[n,r,k,s] envR[c,k,e] v[b,q,r,e],  */
/*   gR[c,k,e] = sum envL[a,m,b] v[a,p,s,c] WL[m,q,n,p] WR[n,r,k,s] v[b,q,r,e],  */
/* each with the structure of the dual of its variable (same sectors and block  */
/* order as envL / envR).  gR is the two-site environment core on the operands, */
/* gL the same core on their mirror image.  The gradient w.r.t. v is           */
/* 2 H_eff v (mpsh_heff_apply).  FP64 and FP32.  stats as for mpsh_envupdate.   */
/* ------------------------------------------------------------------------ */
typedef struct mpsh_heff_adjoint_s* mpsh_heff_adjoint;

mpsh_status mpsh_heff_adjoint_plan(mpsh_struct v, mpsh_struct envL, mpsh_struct envR, mpsh_struct WL,
                                   mpsh_struct WR, mpsh_heff_adjoint* plan, mpsh_struct* gL, mpsh_struct* gR);
size_t mpsh_heff_adjoint_workspace(mpsh_heff_adjoint plan, mpsh_dtype dt);
mpsh_status mpsh_heff_adjoint_stats(mpsh_heff_adjoint plan, int64_t* stats);
mpsh_status mpsh_heff_adjoint_execute(mpsh_heff_adjoint plan, mpsh_dtype dt, const void* v, const void* envL,
                                      const void* envR, const void* WL, const void* WR, void* gL, void* gR,
                                      void* workspace, void* stream);
void mpsh_heff_adjoint_destroy(mpsh_heff_adjoint plan);

/* ------------------------------------------------------------------------ */
/* Batched per-block FP64 SVD: the per-fused-sector SVD of svdtrunc / svd     */
/* (src/mpsutils.jl:116-179; LAPACK gesdd, retried with gesvd on failure,     */
/* src/symtools.jl:14-19).  Block b is the m[b] x n[b] column-major matrix at */
/* A + a_off[b].  Outputs are packed block after block, k = min(m, n):        */
/* S (k values, descending), U (m x k), Vt (k x n), offsets from              */
/* mpsh_svd_sizes (totals = {S, U, Vt} element counts).  Device buffers;     */
/* workspace of mpsh_svd_workspace() bytes, 16-byte aligned.  Synchronizes   */
/* the stream once (convergence flags).                                       */
/* ------------------------------------------------------------------------ */
typedef struct mpsh_svd_s* mpsh_svd;

mpsh_status mpsh_svd_plan(int64_t nblocks, const int64_t* m, const int64_t* n, const int64_t* a_off,
                          mpsh_svd* plan);
mpsh_status mpsh_svd_sizes(mpsh_svd plan, int64_t* k, int64_t* s_off, int64_t* u_off, int64_t* v_off,
                           int64_t* totals);
size_t mpsh_svd_workspace(mpsh_svd plan);
mpsh_status mpsh_svd_execute(mpsh_svd plan, const double* A, double* S, double* U, double* Vt, void* workspace,
                             void* stream);
void mpsh_svd_destroy(mpsh_svd plan);

/* The truncated factors of svdtrunc as packed SymTensors in one launch: the    */
/* kept leading columns of each U block, the kept leading rows of each Vt block */
/* (src/mpsutils.jl:159-161, assembled at :171-176), optionally times the kept  */
/* singular values (u*s, s*v).  FP64 device buffers.  table (device, 8-byte     */
/* aligned) = nblk records of MPSH_GATHER_FIELDS int64                          */
/*   {src_off, rs, cs, rows, cols, dst_off, sc_off, smode}                      */
/* followed by ntiles int64 (block << 32 | tile): every block's                 */
/* ceil(rows*cols / 4096) tiles, 0-based.  For each block and i < rows, j < cols */
/*   dst[dst_off + i + rows*j] = src[src_off + i*rs + j*cs] * f,                */
/*   f = 1 (smode 0), scale[sc_off + i] (1), scale[sc_off + j] (2).             */
/* The table is read on the device only: the caller guarantees that every     */
/* record stays inside src, scale and dst, and passes scale = NULL only when   */
/* every smode is 0.                                                          */
#define MPSH_GATHER_FIELDS 8
mpsh_status mpsh_gather_blocks(int32_t nblk, int64_t ntiles, const int64_t* table, const double* src,
                               const double* scale, double* dst, void* stream);

/* ------------------------------------------------------------------------ */
/* Block-wise vector operations on packed buffers (KrylovKit's interface on   */
/* SymTensors, src/symtensor.jl:266-323).  n = number of elements.            */
/* ------------------------------------------------------------------------ */
#define MPSH_DOT_PARTIALS 1024
/* result (device, one double) = sum x*y; partials: device scratch of MPSH_DOT_PARTIALS doubles */
mpsh_status mpsh_dot(mpsh_dtype dt, int64_t n, const void* x, const void* y, double* result, double* partials,
                     void* stream);
/* y = a*x + b*y   (axpby!, src/symtensor.jl:299-308) */
mpsh_status mpsh_axpby(mpsh_dtype dt, int64_t n, double a, const void* x, double b, void* y, void* stream);
/* x = a*x   (rmul!, src/symtensor.jl:280-285) */
mpsh_status mpsh_scal(mpsh_dtype dt, int64_t n, double a, void* x, void* stream);
/* dst = (dst type) src; dtypes MPSH_F64 / MPSH_F32 */
mpsh_status mpsh_convert(mpsh_dtype src_dt, const void* src, mpsh_dtype dst_dt, void* dst, int64_t n,
                         void* stream);

/* ------------------------------------------------------------------------ */
/* Lanczos basis operations for the eigensolve on the device (the role of    */
/* KrylovKit.eigsolve's orthogonalisation, src/dmrg.jl:77-82).  FP64; basis  */
/* vector i is V + i*ldv (n elements); scalars stay in device memory so one  */
/* Lanczos iteration needs no host round trip and can be graph-captured.     */
/* partials: device scratch of k * MPSH_LZ_PARTS doubles.                    */
/* ------------------------------------------------------------------------ */
#define MPSH_LZ_PARTS 256
/* c[i] = <V_i, w>, i < k */
mpsh_status mpsh_lanczos_dots(int32_t k, int64_t n, const double* V, int64_t ldv, const double* w, double* c,
                              double* partials, void* stream);
/* w += sign * sum_{i<k} c[i] V_i   (c in device memory) */
mpsh_status mpsh_lanczos_update(int32_t k, int64_t n, const double* V, int64_t ldv, const double* c, double sign,
                                double* w, void* stream);
/* *beta = ||w||, out = w / *beta (out may alias w) */
mpsh_status mpsh_lanczos_normalize(int64_t n, const double* w, double* out, double* beta, double* partials,
                                   void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MPSH_H */
