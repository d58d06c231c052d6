/*
 * flexmm_amd -- C-ABI of the MI355X (gfx950) 2D kernel-independent FMM.
 *
 * The reference (dbstein/flexmm, kifmm2d path) exposes a Python class, not an
 * FFI; each entry point below replaces one step of that class, cited as
 * file:line of the reference.  All pointers named x, y, tau, tx, ty, out, dst
 * are DEVICE pointers (for example torch.Tensor.data_ptr() of a CUDA/HIP
 * tensor); `stream` is a hipStream_t (torch.cuda.current_stream().cuda_stream)
 * or NULL for the default stream.  Work is enqueued on that stream; the tree
 * build waits on the host for the few counts it sizes its per-level arrays
 * by, which its kernels hand over through pinned host memory as they are
 * made, so the stream keeps running (it is not drained).
 *
 * Memory: a plan owns its device buffers (tree, expansions, target scratch).
 * They are allocated with hipMalloc when a call first needs them and grow
 * only when a later call needs more (steady-state calls of the same sizes
 * allocate nothing); a buffer that grows first waits for the stream of the
 * call in progress, never for the whole device.  fmm2d_plan_destroy waits
 * for the work of the plan's last call and frees them.  Calls on one plan
 * are ordered by their streams: a call on another stream than the previous
 * one must be ordered after it by the caller (an event), as for any
 * stream-bound handle.  fmm2d_evaluate_noalloc allocates nothing at all.
 *
 * Inputs: source coordinates that are NaN or infinite are refused by
 * fmm2d_tree_build (FMM2D_E_ARG).  A NaN or infinite target coordinate gives
 * non-finite outputs for that target only.  Sources and targets outside the
 * root box are located by the same midpoint comparisons as inside (to an
 * edge cell, as the reference's Tree and locate_points do, tree.py:127-189,
 * 703-727), so targets outside it are extrapolated from that cell's local
 * expansion exactly as in the reference.
 *
 * Return value: 0 on success, < 0 for an argument/state/limit error
 * (FMM2D_E_*), > 0 for a HIP runtime error (hipError_t).  The message of the
 * last failure on the calling thread is returned by fmm2d_last_error().
 * Nothing is thrown across the ABI.  A plan is not thread-safe; distinct
 * plans may be used concurrently.
 */
#ifndef FLEXMM_AMD_H
#define FLEXMM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FMM2D_ABI_VERSION 1

/* error codes (negative) */
#define FMM2D_E_ARG (-1)
#define FMM2D_E_STATE (-2)
#define FMM2D_E_LIMIT (-3)
#define FMM2D_E_INTERNAL (-4)

/* kernels (the reference's user kernel_eval, fmm.py:38; examples/pure_kifmm_laplace.py:24-29) */
#define FMM2D_KERNEL_LAPLACE 0 /* -log(|x-y|^2)/(4 pi) */
#define FMM2D_KERNEL_MODIFIED_HELMHOLTZ 1 /* K0(k |x-y|)/(2 pi), kparams = {k}, k > 0
                                             (examples/pure_kifmm_modified_helmholtz_gradient.py:36-68) */
#define FMM2D_KERNEL_STOKESLET 2 /* 2D Stokeslet: velocity (-log r I + r r^T / r^2) f / (4 pi) and
                                    pressure (r . f) / (2 pi r^2) (kifmm2d/vector/stokes_fmm.py:33-61);
                                    its far field is three harmonic fields carried by Laplace
                                    equivalent densities, so the plan takes the Laplace
                                    operators */

/* evaluators (register_evaluator, helpers.py:176-238) */
#define FMM2D_EVAL_POTENTIAL 0 /* 1 output: u */
#define FMM2D_EVAL_STOKES 2    /* 3 outputs: velocity u, v and pressure p (Stokeslet kernel
                                  only; stokes_fmm.py:827-840) */
#define FMM2D_EVAL_DIPOLE 3    /* 1 output: u of Laplace single + double layer sources
                                  (needs fmm2d_upward_dipole; the evaluator
                                  'slp+dlp->u' of examples/pure_kifmm_laplace_dipoles.py:
                                  139-145); a target on a source without self-exclusion
                                  gives NaN */
#define FMM2D_EVAL_STOKES_DIPOLE 4 /* 3 outputs: u, v, p of Stokeslets plus stresslets (needs
                                      fmm2d_upward_dipole on a Stokeslet plan; replaces
                                      evaluate_to_points_dipole, stokes_fmm.py:844-858);
                                      a target on a source without self-exclusion
                                      gives NaN */
#define FMM2D_EVAL_GRADIENT 1  /* 3 outputs: u, du/dx_t, du/dy_t
                                  (examples/pure_kifmm_laplace_gradient.py:32-55,
                                  pure_kifmm_modified_helmholtz_gradient.py:45-68) */

/* precision */
#define FMM2D_PREC_FP64 0  /* every step in FP64 */
#define FMM2D_PREC_MIXED 1 /* FP32 MFMA M2L (constant mode in FP64; the Stokeslet fields
                              translated in the target box's frame) and, for the Laplace
                              potential, FP32 near field for leaf colleagues at
                              separations >= w/64; FP64 elsewhere.  An evaluation
                              with fewer than 32 targets per leaf on average takes
                              the FP64 near field (faster there, and more accurate).
                              Black-box plans and Stokes plans with stresslets
                              (fmm2d_upward_dipole) keep the FP64 M2L: an FP32
                              accumulation alone costs them 1e-7 to 4e-7 */
#define FMM2D_PREC_MIXED_M2L 2 /* FP32 MFMA M2L only (black-box and stresslet plans:
                                  FP64); near field in FP64 */

/* FP64 M2L form (fmm2d_set_m2l_mode) */
#define FMM2D_M2L_SVD 0   /* the reference's M2L_Evaluator (helpers.py:17-45): a slot of
                             numerical rank r (singular values > 1e-14) with 2 r < p is
                             applied as M2 (M1 q), a null slot adds nothing, others dense */
#define FMM2D_M2L_DENSE 1 /* default: every slot applied as its dense matrix.  On MFMA the
                             compressed form saves little (its ranks, 15-29 at p = 60,
                             pad to 32 rows) and measured slower at every size tried */

/* tree build (fmm2d_set_tree_mode) */
#define FMM2D_TREE_AUTO 0     /* default: points that form a complete quadtree (every cell of
                                 depth D-1 holds more than Ncutoff points, every cell of depth D
                                 between 1 and Ncutoff, 2 <= D < 10) take the closed-form build
                                 of that tree; all others the general build */
#define FMM2D_TREE_GENERAL 1  /* always the general build (sort, refinement, 2:1 balance,
                                 X-list); the two give identical trees */

/* tree arrays for fmm2d_tree_export */
#define FMM2D_TREE_KEYS 0       /* u64[n]   Morton keys of the sorted sources            */
#define FMM2D_TREE_PERM 1       /* u32[n]   tree position -> input index (Tree.ordv)      */
#define FMM2D_TREE_SX 2         /* f64[n]   sorted x (Tree.x)                             */
#define FMM2D_TREE_SY 3         /* f64[n]   sorted y (Tree.y)                             */
#define FMM2D_TREE_CODE 4       /* u64[NN]  Morton code of each node in its level          */
#define FMM2D_TREE_LEVEL 5      /* i32[NN]                                                */
#define FMM2D_TREE_BOT 6        /* i32[NN]  Level.bot_ind                                 */
#define FMM2D_TREE_TOP 7        /* i32[NN]  Level.top_ind                                 */
#define FMM2D_TREE_PARENT 8     /* i32[NN]  global parent index                           */
#define FMM2D_TREE_CHILD 9      /* i32[NN]  global first child or -1 (Level.children_ind) */
#define FMM2D_TREE_FLAGS 10     /* i32[NN]  bits: 1 leaf, 2 fake_leaf, 4 Xlist,
                                   8 compute_upwards, 16 compute_downwards,
                                   32 has children, 64 has sources                       */
#define FMM2D_TREE_COLLEAGUES 11 /* i32[NN*9] colleagues, slot (dx+1)*3+(dy+1)           */
#define FMM2D_TREE_XMID 12      /* f64[NN]                                                */
#define FMM2D_TREE_YMID 13      /* f64[NN]                                                */
#define FMM2D_TREE_ROW 14       /* i32[NN]  multipole row (this_density_ind) or -1        */
#define FMM2D_TREE_M2L 15       /* i32[NN*40] multipole rows of the interaction list     */
#define FMM2D_TREE_EVALID 16    /* i32[NN]  index among leaves or -1                      */

/* expansion arrays for fmm2d_export_expansions (rows of PD doubles) */
#define FMM2D_EXP_CHECK 0       /* S2M check potentials (upward partial multipoles)   */
#define FMM2D_EXP_MULTIPOLE 1   /* equivalent densities (FMM.multipoles)              */
#define FMM2D_EXP_LOCAL 2       /* local equivalent densities (FMM.Local_Expansions)  */
#define FMM2D_EXP_TAU_SORTED 3  /* charges in tree order (FMM.tau_ordered)            */

/* phases timed by fmm2d_set_profiling / fmm2d_phase_times */
#define FMM2D_PHASE_TREE 0        /* fmm2d_tree_build                         */
#define FMM2D_PHASE_UPWARD 1      /* fmm2d_upward                             */
#define FMM2D_PHASE_DOWNWARD 2    /* fmm2d_downward                           */
#define FMM2D_PHASE_EVALUATE 3    /* fmm2d_evaluate (locate, sort, L2T + P2P) */
#define FMM2D_PHASE_EVAL_KERNEL 4 /* the fused L2T + P2P kernel alone         */
#define FMM2D_PHASE_S2M 5         /* the S2M kernel alone                     */
#define FMM2D_PHASE_M2L 6         /* the M2L kernel(s) of the downward pass   */
#define FMM2D_NUM_PHASES 7

int fmm2d_abi_version(void);
const char* fmm2d_last_error(void);

/* Create/destroy a plan: one FMM for one kernel and Nequiv = p (4 <= p <= 64).
 * Replaces FMM.__init__'s helper lookup (fmm.py:38-43, 61-65). */
int fmm2d_plan_create(int kernel_id, const double* kparams, int nkparams, int p, int precision,
                      void** plan);
/* Create a black-box (Chebyshev interpolation) plan of the given order
 * (2 <= order <= 8; Nequiv = order^2 Chebyshev moments / coefficients per
 * box) for the Laplace or modified Helmholtz kernel.  Replaces
 * BB_FMM.__init__ and BB_Helper.prepare (flexmm2d/bbfmm.py:122-209): the
 * level operators passed to fmm2d_set_operators are M2MC / L2LC
 * (bbfmm.py:138-157) and the 40 M2L matrices VI K VI^T (:97-117), with
 * identity pseudo-inverse blocks and the level's 2/width as the first
 * circle entry.  fmm2d_upward takes charges (S2M = Chebyshev moments,
 * :173-184); fmm2d_evaluate takes FMM2D_EVAL_POTENTIAL only (L2T = the
 * Chebyshev sum, :185-199).  FMM2D_PREC_MIXED (FP32 near field) needs the
 * Laplace kernel. */
int fmm2d_plan_create_cheb(int kernel_id, const double* kparams, int nkparams, int order,
                           int precision, void** plan);
int fmm2d_plan_destroy(void* plan);

/* Build the quadtree of n sources in the square box bbox = [xmin, xmax, ymin,
 * ymax] (width xmax - xmin) with leaf capacity ncut.
 * Replaces Tree(x, y, Ncutoff, bbox) (fmm.py:68-72, tree.py:451-518). */
int fmm2d_tree_build(void* plan, const double* x, const double* y, int64_t n, const double* bbox,
                     int64_t ncut, void* stream);

/* Choose the tree build (FMM2D_TREE_*) of later fmm2d_tree_build calls. */
int fmm2d_set_tree_mode(void* plan, int mode);

/* Host summary: out[0..7] = levels L, nodes, multipole rows, leaves, S2M boxes,
 * sources, row stride PD, ncut; then per level d: out[8+3d] node offset,
 * out[9+3d] node count, out[10+3d] first multipole row; out[8+3L] = rows.
 * Needs nout >= 9 + 3L.  Replaces the Tree.Levels attributes (tree.py:275-308). */
int fmm2d_tree_info(void* plan, int64_t* out, int64_t nout);

/* Copy one tree array (FMM2D_TREE_*) into a device buffer of `bytes` bytes. */
int fmm2d_tree_export(void* plan, int which, void* dst, int64_t bytes, void* stream);

/* Choose how the FP64 M2L applies the translation matrices (FMM2D_M2L_*);
 * takes effect at the next fmm2d_downward.  The FP32 M2L of the mixed forms
 * always applies the dense residuals. */
int fmm2d_set_m2l_mode(void* plan, int mode);

/* Doubles per level of the operator block for Nequiv = p, or -1. */
int64_t fmm2d_operator_size(int p);

/* Set the precomputed FP64 operators of one tree level from a device array of
 * fmm2d_operator_size(p) doubles (copied; the caller may free it afterwards):
 * 50 blocks of [PD][PD] -- diag(1/s)U^T, V, the 4 M2M column blocks of M2MC,
 * their 4 transposes (L2L), the 40 M2L matrices in slot order -- then the
 * large-circle x and y (PD each), then for the mixed-precision M2L the row
 * means c_s of each M2L matrix ([40][PD] doubles) and the FP32 residuals
 * R_s = M_s - c_s 1^T (40 blocks of PD^2 floats, i.e. 20 PD^2 doubles of
 * storage, each in MFMA operand order: entry ((rb PD/16 + mp) 64 + lane) 4 + e
 * holds R[16 rb + lane%16][8 m + 2 (lane/16) + e%2], m = 2 mp + e/2), then
 * the SVD-compressed M2L of M2L_Evaluator (helpers.py:17-45): 64 slot ranks
 * as doubles (r > 0 compressed, 0 null, -1 dense; slots 40..63 unused) and 40
 * [PD][PD] blocks, rows 0 .. PD/2-1 holding M1 = diag(s_:r) V_:r^T and rows
 * PD/2 .. PD-1 holding M2 = U_:r two rows per block row
 * (block[PD/2 + i/2][(i%2) PD/2 + j] = M2[i][j]).
 * PD = 32 for p <= 32, else 64.  Replaces Precomputation (helpers.py:250-317). */
int fmm2d_set_operators(void* plan, int level, const double* dev_ops, int64_t n, void* stream);
int fmm2d_clear_operators(void* plan);

/* Upward pass from charges tau (input order, n values; for the Stokeslet
 * kernel 2n values: the forces f_x[n] then f_y[n]).
 * Replaces FMM.upwards_pass (fmm.py:73-124; stokes_fmm.py:689-718). */
int fmm2d_upward(void* plan, const double* tau, int64_t n, void* stream);

/* Upward pass for Laplace sources carrying a single and a double layer
 * density: tau = 2n values (charges, then dipole strengths), dvec = 2n values
 * (dipole directions: x components, then y components), both in input order.
 * Source kernel -log(d^2)/(4 pi) tau_0 + (n.d)/d^2/(2 pi) tau_1, d = t - s.
 * Replaces FMM.upwards_pass with a registered upwardor and extras
 * (fmm.py:73-124 with extras, helpers.py:155-174; the 'slp+dlp' upwardor of
 * examples/pure_kifmm_laplace_dipoles.py:32-40,139-141).  Evaluate with
 * FMM2D_EVAL_DIPOLE.
 * On a Stokeslet plan: tau = 4n values (forces f_x, f_y, then double-layer
 * densities sigma_x, sigma_y), dvec = 2n values (normals n_x, then n_y);
 * stresslet velocity (1/pi)(r.n)(r.sigma) r/|r|^4 and pressure
 * (1/pi)[-(n.sigma)/|r|^2 + 2(r.n)(r.sigma)/|r|^4], r = t - s.  Replaces
 * FMM.build_expansions_dipole (stokes_fmm.py:738-826; kernels :129-235).
 * Evaluate with FMM2D_EVAL_STOKES_DIPOLE. */
int fmm2d_upward_dipole(void* plan, const double* tau, const double* dvec, int64_t n,
                        void* stream);

/* Downward pass.  Replaces FMM.downwards_pass (fmm.py:125-166). */
int fmm2d_downward(void* plan, void* stream);

/* Evaluate at nt targets into out[outputs][nt] (input target order).
 * check != 0 skips sources that coincide exactly with the target.  nt = 0
 * does nothing (tx, ty and out may then be null).
 * Replaces FMM.evaluate_to_points (fmm.py:182-199). */
int fmm2d_evaluate(void* plan, const double* tx, const double* ty, int64_t nt, int evaluator,
                   int check, double* out, void* stream);

/* fmm2d_evaluate without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as fmm2d_evaluate; the target-side scratch
 * (target boxes, counts, leaf-ordered targets, work items) lives in one
 * device array of the library (about 38 MB, held on each GPU the library runs
 * on until the process ends), so:
 *   - nt <= 1048576 targets and trees of <= 262144 evaluation boxes (leaves
 *     and X-list parents), else FMM2D_E_LIMIT;
 *   - every call on one GPU shares that GPU's array: calls that overlap on
 *     one GPU (other streams, other threads) race;
 *   - the plan's device must be current; a mixed-precision plan needs
 *     fmm2d_downward to have run (it prepares the FP32 near-field sources);
 *   - phase timing (fmm2d_set_profiling) does not time these calls.
 * Zeroing uses a fill kernel (no memset node). */
int fmm2d_evaluate_noalloc(void* plan, const double* tx, const double* ty, int64_t nt,
                           int evaluator, int check, double* out, void* stream);

/* Copy an expansion array (FMM2D_EXP_*) into a device buffer of nvals doubles. */
int fmm2d_export_expansions(void* plan, int which, double* dst, int64_t nvals, void* stream);

/* Phase timing with HIP events recorded on the caller's stream (off by
 * default): every run of a phase records its own pair of events.
 * fmm2d_phase_times waits for them and writes, for each FMM2D_PHASE_*, the
 * mean duration in milliseconds of the runs recorded since profiling was
 * enabled or since the previous fmm2d_phase_times call (-1 if none). */
int fmm2d_set_profiling(void* plan, int enable);

/* Algorithmic work of the current tree and the last evaluation (synchronous):
 * out[0] P2P pairs, [1] targets with an L2T (leaf level >= 2), [2] targets,
 * [3] M2L box pairs, [4] S2M source points, [5] multipole rows (upward
 * pseudo-inverse applies), [6] local expansions (downward pseudo-inverse
 * applies), [7] M2M parents, [8] L2L children.  Needs n >= 9. */
int fmm2d_work_counts(void* plan, int64_t* out, int n);
int fmm2d_phase_times(void* plan, double* ms, int n);

#ifdef __cplusplus
}
#endif

#endif /* FLEXMM_AMD_H */
