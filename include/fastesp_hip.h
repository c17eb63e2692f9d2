/*
 * fastesp_hip.h -- C ABI of libfastesp_hip.so, the MI355X (gfx950) backend of
 * fastesp's ESP hot path.
 *
 * The reference exposes this path as two processes exchanging files
 * (fastesp/cli.py:143-163): the table builder `build_pairs <basis.bin>
 * <pairs.bin> [tol]` (fastesp/native/build_pairs.cpp:162-168) and the CUDA
 * evaluator `esp_gpu <pairs.bin> <out.txt> [--repeats N]`
 * (fastesp/native/esp_gpu.cu:232-242), wrapped in Python by
 * `build_table(...)` / `PairTable.esp()` (fastesp/table.py:73-104,176-259).
 * Each entry point below names the reference interface it replaces.
 *
 * Conventions
 *  - Every function returns 0 (FE_OK) on success and a negative FE_ERR_* code
 *    on failure; it never calls exit().  fe_last_error() returns a
 *    thread-local message describing the last failure of the calling thread.
 *  - Host arrays passed in are borrowed (C-contiguous, little-endian f64 /
 *    i32, caller-owned) and not retained after the call returns.
 *  - `ptr_kind` says where point/output buffers live: FE_PTR_HOST (the library
 *    copies through device staging buffers and synchronises the stream before
 *    returning) or FE_PTR_DEVICE (device pointers; work is only enqueued on
 *    `stream`, which is a hipStream_t passed as void*, NULL = default stream).
 *  - One fe_ctx per device.  The context's entry points take a lock, so
 *    calls from several host threads queue up; but the context's workspace is
 *    shared by every stream, so device-pointer calls that enqueue work on
 *    different streams must be ordered by the caller (host-pointer calls
 *    synchronise their stream before returning and are safe).
 *  - Units: Bohr in, Hartree/e out (fastesp/table.py:73-104).
 */
#ifndef FASTESP_HIP_H
#define FASTESP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FE_OK 0
#define FE_ERR_ARG (-1)         /* invalid argument / shape                    */
#define FE_ERR_HIP (-2)         /* HIP runtime error                           */
#define FE_ERR_UNSUPPORTED (-3) /* e.g. L = l_A + l_B above FE_LMAX            */
#define FE_ERR_FORMAT (-4)      /* table not in the canonical per-L layout     */
#define FE_ERR_ALLOC (-5)       /* device or host allocation failed            */
#define FE_ERR_STATE (-6)       /* wrong object / use after free               */

#define FE_PTR_HOST 0
#define FE_PTR_DEVICE 1

#define FE_PREC_FP64 0          /* FP64 throughout                             */
#define FE_PREC_MIXED 1         /* FP32 monomial products, FP64 Boys + sums    */

#define FE_PART_NUCLEAR 1       /* + sum_A Z_A/|C-R_A|   (fastesp/table.py:95-100) */
#define FE_PART_ELECTRON 2      /* - sum_pairs K sum_k g R^a F_m (table.py:73-93)  */
#define FE_PART_ESP 3           /* nuclear - electron      (table.py:102-104)     */

typedef struct fe_ctx fe_ctx;
typedef struct fe_table fe_table;
typedef struct fe_density fe_density;

/* Library version (major*10000 + minor*100 + patch). */
int fe_version(void);

/* Thread-local description of the last error ("" if none). */
const char* fe_last_error(void);

/* Create a context on HIP device `device`: uploads the Boys tables and
 * caches device properties.  Replaces the per-process CUDA context set-up and
 * warm-up launch of esp_gpu (fastesp/native/esp_gpu.cu:244-297). */
int fe_init(int device, fe_ctx** ctx);

/* Free every device buffer of the context (tables must be freed first). */
int fe_finalize(fe_ctx* ctx);

/* Release the context's workspace (the scratch that table builds and
 * evaluations grow on demand and keep for the next call: about 0.1 GB after a
 * 1634-function build, 0.7 GB after a 4538-function one) and the cached
 * basis set-up.  Waits for all work on the device first.  The next call
 * allocates again what it needs.  No counterpart in the reference, whose
 * processes free everything at exit (fastesp/cli.py:143-163). */
int fe_ctx_trim(fe_ctx* ctx);

/* Upload an existing flat pair table (the pairs.bin fields of
 * fastesp/table.py:11-24, as written by build_pairs.cpp:386-421 or
 * PairTable.write, fastesp/table.py:119-141) and repack it into the per-L
 * device layout.  Replaces the pairs.bin loader + H2D copies of esp_gpu
 * (fastesp/native/esp_gpu.cu:96-134,251-277).  Every pair must use the
 * canonical term list and monomial schedule of its class (FE_ERR_FORMAT
 * otherwise).  Z/xyz describe the nuclei (n_atoms may be 0). */
int fe_table_from_pairs(fe_ctx* ctx, int n_pairs, int n_terms,
                        const double* pp_exp, const double* pp_cx,
                        const double* pp_cy, const double* pp_cz,
                        const double* pp_K, const int32_t* pp_nt,
                        const int32_t* pp_off, const int32_t* pp_nmono,
                        const int32_t* pp_monooff, const int32_t* mo_parent,
                        const int32_t* mo_axis, const int32_t* tm_aidx,
                        const int32_t* tm_m, const double* tm_g, int n_atoms,
                        const double* Z, const double* xyz, fe_table** out);

/* Build the pair table on the device from a Cartesian basis, a density
 * matrix and the ESP points (used only for the screen).  Arguments are the
 * basis.bin fields of fastesp/table.py:268-316 (one entry per Cartesian
 * component, primitives duplicated per component, components of a shell
 * contiguous: every component of the shell's l exactly once, in cart_order or
 * any other order; a shell that repeats or lacks a component is
 * FE_ERR_FORMAT).  Replaces `build_pairs <basis.bin> <pairs.bin>
 * [tol]` (fastesp/native/build_pairs.cpp:162-425) and the NumPy builder
 * ESPEngine + build_table (fastesp/esp.py:168-257, fastesp/table.py:176-259).
 * rule 0: the NumPy screen, keep = !(K bound / max(dmin, 0.05) < tol)
 *         (fastesp/table.py:190-194);
 * rule 1: the C++ screen, keep = !(K bound / max(dmin - reff, 0.05) < tol) and
 *         G not all zero and c_i c_j != 0 (fastesp/native/build_pairs.cpp:
 *         243,309,318-328);
 * rule 2: the ESPEngine screen, keep = !(K bound / max(dmax - reff, 0.05) < tol)
 *         with dmax the largest distance to any point (fastesp/esp.py:265-270).
 * `ptr_kind` applies to dens (n_func x n_func, row-major) and pts (n_points x 3);
 * the basis arrays are host arrays.  Shells up to f (L = l_A + l_B <= 6). */
int fe_table_build(fe_ctx* ctx, int n_func, int n_prim, const int32_t* fn_atom,
                   const int32_t* fn_nprim, const int32_t* fn_off,
                   const int32_t* fn_lx, const int32_t* fn_ly, const int32_t* fn_lz,
                   const int32_t* fn_shell, const double* fc, const double* pexp,
                   const double* pcoef, const double* dens, int n_points,
                   const double* pts, int n_atoms, const double* Z,
                   const double* xyz, double tol, int rule, int ptr_kind,
                   void* stream, fe_table** out);

/* Build a table made by fe_table_build again, in place, for a new density
 * and/or new points (same basis, same rule semantics as fe_table_build).  No
 * counterpart in the reference, which rebuilds from scratch
 * (fastesp/cli.py:143-163); it exists so that a build can run where a host
 * round trip or an allocation cannot: in a differentiable training loop or
 * under HIP graph capture.
 *  - The first call resizes the table to hold every primitive pair with
 *    K >= 1e-30 (a bound for any density and points), which allocates and
 *    synchronises; later calls allocate nothing and never synchronise with
 *    the host (device pointers), so they can be captured.
 *  - The context's build workspace is reused; it grows (allocates) only when
 *    n_points exceeds every earlier build's.
 *  - The new counts stay on the device: evaluations read them there, and
 *    host queries (fe_table_info, fe_table_export, fe_table_stats, a dL/dG
 *    output of fe_esp_backward, fe_batch_create) wait for the rebuild and read
 *    them back.  A batch made from the table must not outlive a rebuild. */
int fe_table_rebuild(fe_table* t, const double* dens, int n_points, const double* pts,
                     double tol, int rule, int ptr_kind, void* stream);

/* Floating-point statistics of a table: out[0] the sum rule
 * sum_pairs K (pi/p)^{3/2} etab_0 = Tr(P S) over all candidate primitive
 * pairs (printed by build_pairs.cpp:423), out[1] kept pairs, out[2] basis
 * functions, out[3] shell pairs (zeros for uploaded tables). */
int fe_table_stats(const fe_table* t, double* out, int n);

/* Copy a table back to the host in the flat pairs.bin layout
 * (fastesp/table.py:11-24), pairs in reference order (shell pairs si <= sj,
 * primitives i-major; fastesp/esp.py:168-193, fastesp/table.py:223-259).
 * Array sizes: n_pairs, n_terms (fe_table_info) and n_mono = sum of pp_nmono
 * (= sum over pairs of the class monomial count).  pp_kbound (K * bound) and
 * pp_reff may be NULL.  Replaces the pairs.bin writer of build_pairs.cpp:386-421. */
int fe_table_export(const fe_table* t, double* pp_exp, double* pp_cx, double* pp_cy,
                    double* pp_cz, double* pp_K, int32_t* pp_nt, int32_t* pp_off,
                    int32_t* pp_nmono, int32_t* pp_monooff, int32_t* mo_parent,
                    int32_t* mo_axis, int32_t* tm_aidx, int32_t* tm_m, double* tm_g,
                    double* pp_kbound, double* pp_reff);

/* Release a table and its device memory. */
int fe_table_free(fe_table* t);

/* Table statistics: info[0] n_pairs, info[1] n_terms, info[2] n_atoms,
 * info[3..3+FE_LMAX] pairs per class L, info[10] device bytes,
 * info[11] candidate pairs before the point screen: primitive pairs with
 * K >= 1e-30 in shell pairs whose density block reaches 1e-15, the
 * reference's primitive-pair list (0 for tables uploaded from pairs.bin). */
int fe_table_info(const fe_table* t, int64_t* info, int n_info);

/* ESP at n points: V = [nuclear] - [electron] as selected by `parts`.
 * pts is (n,3) f64; V is (n,) f64.  Replaces esp_pairs_kernel +
 * esp_reduce_kernel (fastesp/native/esp_gpu.cu:161-229) and
 * PairTable.esp/electron/nuclear (fastesp/table.py:73-104).
 * `precision` is FE_PREC_FP64 or FE_PREC_MIXED. */
int fe_esp_forward(fe_table* t, const double* pts, int64_t n, int precision,
                   int parts, int ptr_kind, void* stream, double* V);

/* V and the field dV/dC (n x 3, f64) in one pass.  No counterpart in the
 * reference (it has no gradients, fastesp/table.py:73-104 is value-only);
 * V is the same quantity as fe_esp_forward(parts = FE_PART_ESP). */
int fe_esp_forward_grad(fe_table* t, const double* pts, int64_t n, int precision,
                        int ptr_kind, void* stream, double* V, double* dVdC);

/* Adjoint of V for a loss with dL/dV = gV (n values): dL/dG_k =
 * -sum_i gV_i K R_i^a F_m (n_terms values, aligned with the exported tm_g,
 * NULL to skip) and dL/dP (n_func x n_func, both triangles filled because the
 * table folds P_ij + P_ji^T, fastesp/esp.py:185-187; NULL to skip; needs a
 * table from fe_table_build).  No counterpart in the reference. */
int fe_esp_backward(fe_table* t, const double* pts, int64_t n, const double* gV,
                    int precision, int ptr_kind, void* stream, double* dLdG,
                    double* dLdP);

/* fe_esp_forward on device buffers that allocates nothing: no hipMalloc,
 * hipFree or allocating library call, so it can run where allocation is not
 * allowed (under HIP graph capture, in tight serving loops, in harnesses that
 * forbid it).  No counterpart in the reference (esp_gpu allocates per run,
 * fastesp/native/esp_gpu.cu:251-277); it is for those callers.
 *  - pts (n x 3) and V (n) are device pointers; work is enqueued on `stream`.
 *  - n <= FE_NOALLOC_MAX_POINTS, else FE_ERR_ARG.
 *  - Its scratch (point sort + partial sums, FE_NOALLOC_SCRATCH_MIB MiB) is a
 *    fixed device array that every GPU holds for the life of the process.
 *    Every call on one GPU shares that GPU's array and keeps the last call's
 *    contents: calls that overlap on one GPU (two streams, or two threads with
 *    different contexts of the same device) race and must be ordered by the
 *    caller.
 *  - FE_PREC_MIXED needs the table's FP32 coefficients, which the first
 *    fe_esp_forward with FE_PREC_MIXED builds (FE_ERR_STATE otherwise).
 *  - Same parts, precision and results as fe_esp_forward; the pair-chunk
 *    count may be lower (fewer partial rows), which can change the last bits. */
#define FE_NOALLOC_MAX_POINTS 65536
#define FE_NOALLOC_SCRATCH_MIB 32
int fe_esp_forward_noalloc(fe_table* t, const double* pts, int64_t n, int precision,
                           int parts, void* stream, double* V);

/* Boys functions F_0..F_M(T) for n arguments with the device algorithm
 * (F is (n, M+1) f64, row-major).  Replaces boys_all
 * (fastesp/native/esp_gpu.cu:36-66) / fastesp.ints.boys (fastesp/ints.py:35-40)
 * for parity testing of the device scheme.  0 <= M <= FE_BOYS_MMAX (8). */
int fe_boys_eval(fe_ctx* ctx, const double* T, int64_t n, int M, int ptr_kind,
                 void* stream, double* F);

/* Tuning knobs of the single-table evaluation (autotuner output).
 * key 0: points per lane of every class (1, 2 or 4); 10 + L: of class L only
 *        (both set the value and the value + field kernels; defaults 2, with
 *        4 for class 6 of the value kernels and 1 for classes 5 and 6 with the
 *        field);
 * key 1: target workgroups per class launch (sets the pair-chunk count;
 *        default 32 per compute unit);
 * key 3: Morton-order the points before evaluation (0/1, default 1);
 * key 4: FP32 routing threshold of FE_PREC_MIXED in units of 1e-12 a.u.
 *        (default 1e8, i.e. 1e-4 a.u.; 0 sends every pair to FP64);
 * key 5: order the pairs of each class by exponent bucket, then Morton code
 *        of P, at table creation (0/1, default 1; export keeps reference order);
 * key 6: streams of the dL/dG class launches of fe_esp_backward and of the
 *        batched backward (the caller's and streams the context owns, joined
 *        before the call returns; largest class first, each to the least
 *        loaded stream): 0 one stream, 1 or 2 two (default), 3 three;
 * key 7: run the evaluation class launches on two streams, each with its own
 *        partial rows: 0 never; 1 always, in two balanced halves; 2 (default)
 *        the value kernels in two balanced halves (config 1: forward 4.31 ->
 *        4.16 ms) and the value + field kernels with the small classes 5 and
 *        6 beside the first large ones;
 * key 20 + L: class L takes the FP32 contraction under FE_PREC_MIXED (0/1;
 *        default 1 for L >= 3, else the class runs the FP64 kernel). */
int fe_set_tuning(fe_ctx* ctx, int key, int64_t value);

/* Phase timing with HIP events on the launch stream (no counterpart in the
 * reference, whose esp_gpu times whole runs, fastesp/native/esp_gpu.cu:299-313).
 * op 1 starts recording; op 0 stops and returns, per phase id < n_phase, the
 * summed milliseconds and the number of launches.  Phase ids: 0 point sort,
 * 1 per-point reduction, 2 gV packing, 3 dL/dP, 8+L evaluation of class L,
 * 16+L dL/dG of class L, 24+L dL/dG chunk reduction of class L.  Launches
 * recorded into a hipGraph are not timed. */
int fe_profile(fe_ctx* ctx, int op, double* ms, int64_t* count, int n_phase);

/* Work counts of the launches fe_esp_forward(_grad) and fe_esp_backward would
 * make for these points with the context's current tuning (no counterpart in
 * the reference; it prices the performed work of the roofline): counts[L] and
 * counts[7 + L] are the (point, pair) evaluations of class L that the
 * evaluation launch (value + field if grad != 0, else value only) takes on the
 * far-field and on the tabulated path, counts[14 + L] and counts[21 + L] the
 * same for the dL/dG launch.  Both launches decide the path per wavefront,
 * and the counts follow the launches' decompositions exactly.  28 values;
 * synchronises the stream. */
int fe_eval_work(fe_table* t, const double* pts, int64_t n, int grad, int ptr_kind, void* stream,
                 int64_t* counts);

/* ---- batched conformers ------------------------------------------------
 * No counterpart in the reference (it evaluates one table per esp_gpu
 * process, fastesp/cli.py:143-163).  A batch references n_conf tables of one
 * context (they must outlive it) and owns the concatenated points (conformer
 * c holds rows [sum_{c'<c} n_pts[c'], ...) of pts, caller order), the outputs
 * and a hipGraph of one step.  Outputs are concatenated the same way; dL/dP
 * of conformer c is its n_func x n_func block after those of 0..c-1. */
typedef struct fe_batch fe_batch;

#define FE_BATCH_FIELD 1   /* forward also computes dV/dC                 */
#define FE_BATCH_DLDP 2    /* step runs loss + backward to dL/dP          */
#define FE_BATCH_REBUILD 4 /* step first builds every table again from the
                              densities of fe_batch_set_density and the
                              batch's current points                      */

#define FE_BOUT_V 0        /* V, n_tot                                    */
#define FE_BOUT_FIELD 1    /* dV/dC, n_tot x 3                            */
#define FE_BOUT_DLDP 2     /* dL/dP blocks, sum_c n_func_c^2              */
#define FE_BOUT_LOSS 3     /* sum_i w_i (V_i - T_i)^2 per conformer       */
#define FE_BOUT_GV 4       /* dL/dV, n_tot                                */

#define FE_BT_ITEMS_PER_WG 1    /* forward work items per resident workgroup */
#define FE_BT_DG_ITEMS_PER_WG 2 /* dL/dG work items per resident workgroup   */
#define FE_BT_CLASS_MASK 3      /* bit L: run class L (timing only)          */
#define FE_BT_PPL0 10           /* + L: points per lane of class L (1, 2, 4) */
#define FE_BT_MIXED0 20         /* + L: class L uses FP32 under FE_PREC_MIXED */

int fe_batch_create(fe_ctx* ctx, int n_conf, fe_table* const* tables, const double* pts,
                    const int64_t* n_pts, int ptr_kind, void* stream, fe_batch** out);
int fe_batch_free(fe_batch* b);
/* new coordinates (same counts); the Morton order is recomputed */
int fe_batch_set_points(fe_batch* b, const double* pts, int ptr_kind, void* stream);
/* loss weights w and targets T (n_tot each; NULL keeps; default w = 1, T = 0) */
int fe_batch_set_loss(fe_batch* b, const double* w, const double* target, int ptr_kind,
                      void* stream);
/* The densities of every conformer (the n_func_c x n_func_c row-major blocks,
 * concatenated in conformer order) and the screen (tol, rule as in
 * fe_table_build) for steps with FE_BATCH_REBUILD, which build every table
 * again on the device before the forward, so a recorded step runs build ->
 * forward -> loss -> backward.  No counterpart in the reference (each of its
 * runs builds one table, build_pairs.cpp:162-425).  The first call rebuilds
 * every table once (fe_table_rebuild: their capacity layout, which allocates
 * and synchronises) and sizes a build workspace owned by the batch; later
 * calls only copy the densities.  Every table must come from fe_table_build;
 * from the first call on the tables are the batch's to rebuild (host queries
 * of a table read its latest counts). */
int fe_batch_set_density(fe_batch* b, const double* dens, double tol, int rule, int ptr_kind,
                         void* stream);
int fe_batch_set_tuning(fe_batch* b, int key, int64_t value);
/* V (and dV/dC with FE_BATCH_FIELD) of every conformer */
int fe_batch_forward(fe_batch* b, int precision, int flags, void* stream);
/* loss per conformer and gV = 2 w (V - T) from the last forward */
int fe_batch_loss(fe_batch* b, void* stream);
/* dL/dP of every conformer for gV (n_tot values; NULL: the loss gradient) */
int fe_batch_backward(fe_batch* b, const double* gV, int ptr_kind, void* stream);
/* forward + loss (+ backward with FE_BATCH_DLDP), eager */
int fe_batch_step(fe_batch* b, int precision, int flags, void* stream);
/* record fe_batch_step(precision, flags) as a hipGraph; fe_batch_launch replays
 * it on any stream.  Tuning changes drop the graph. */
int fe_batch_capture(fe_batch* b, int precision, int flags);
int fe_batch_launch(fe_batch* b, void* stream);
int fe_batch_read(fe_batch* b, int what, int ptr_kind, void* stream, double* dst);
/* [n_conf, n_points, n_pairs, forward items, dG items, partial bytes,
 *  dL/dP doubles, shell pairs, captured, items/WG, dG items/WG, class mask] */
int fe_batch_info(const fe_batch* b, int64_t* out, int n_out);

/* ---- electron density and surface points ------------------------------
 * rho(C) = sum_{mu nu} phi_mu(C) P_{mu nu} phi_nu(C): replaces `our_density`
 * of benchmarks/check_density_mwfn.py:60-75 (the per-point density that shares
 * the ESP's basis).  fe_density_create takes the basis.bin fields of
 * fastesp/table.py:268-293 (host arrays) and the density (n_func x n_func,
 * row-major; host or device per ptr_kind; only its symmetric part enters);
 * fe_density_eval writes rho at n_points points (host or device buffers).
 * fe_density_info: [n_func, padded n_func, matrix-core FLOPs of the last
 * evaluation, chunk capacity]. */
int fe_density_create(fe_ctx* ctx, int n_func, int n_prim, const int32_t* fn_nprim,
                      const int32_t* fn_off, const int32_t* fn_lx, const int32_t* fn_ly,
                      const int32_t* fn_lz, const double* fc, const double* pexp,
                      const double* pcoef, const double* dens, int ptr_kind, void* stream,
                      fe_density** out);
int fe_density_eval(fe_density* d, const double* pts, int64_t n_points, int ptr_kind,
                    void* stream, double* rho);
int fe_density_info(const fe_density* d, int64_t* info, int n_info);
/* key 0: 1 = 128 x 64 point x function tiles, 0 = 64 x 64 (default) */
int fe_density_set_tuning(fe_density* d, int key, int64_t value);
int fe_density_free(fe_density* d);

/* Candidate points of the vdW surface: counts[a] Fibonacci-sphere points of
 * radius rad[a] about atom a (Bohr), atom-major, and ok[k] = 1 when point k
 * lies outside every sphere (|p - R_b| >= rad_b - 1e-9).  The device part of
 * examples/vdw_surface_esp.py:38-52 (vdw_surface_points); host arrays. */
int fe_vdw_candidates(fe_ctx* ctx, int n_atoms, const double* xyz, const double* rad,
                      const int64_t* counts, double* pts, int32_t* ok);

#ifdef __cplusplus
}
#endif
#endif /* FASTESP_HIP_H */
