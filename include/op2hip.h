/* op2hip: C ABI of the MI355X (gfx950) backend of the op2 programming model.
 *
 * Library: pyop2_amd/lib/libop2hip.so (built by `python -m pyop2_amd.build`).
 * Python binds it with ctypes (pyop2_amd/_native.py); any FFI can.
 *
 * Conventions
 *   - every call returns int status: 0 ok, < 0 error; op2hip_last_error()
 *     returns the thread-local message of the last failure.  No C++
 *     exception crosses the ABI.
 *   - device buffers are passed as plain pointers and are owned by the
 *     caller; opaque handles (modules, plans, sparsity builds) have an
 *     explicit destroy call.
 *   - every call that enqueues device work takes a hipStream_t as void*.
 *   - one host thread per handle; handles are not thread-safe.
 *
 * Each declaration names the reference interface it replaces
 * (paths relative to the reference repository gbts/PyOP2).
 */
#ifndef OP2HIP_H
#define OP2HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime --------------------------------------------------------------- */

/* Message of the last failing call on this thread. */
const char* op2hip_last_error(void);

/* Select and initialise GPU `device`.
 * Replaces: op_init of OP2-Common, pyop2/op_lib_core.pyx:131-154 (called
 * from op2.init, pyop2/op2.py:82). */
int op2hip_init(int device);

/* LDS bytes per workgroup and per compute unit, compute-unit count. */
int op2hip_device_limits(int device, int* lds_per_block, int* num_cu, int* lds_per_cu);

/* gcnArchName of `device` into `out`. */
int op2hip_device_arch(int device, char* out, int outlen);

/* ---- JIT of generated loops ---------------------------------------------------
 * Replaces the compile step of a JITModule: instant.inline_with_numpy in
 * pyop2/host.py:225-261 and pycuda SourceModule in pyop2/cuda.py:602-651.
 * op2hip_compile needs no GPU; it keeps an on-disk code-object cache in
 * `cache_dir` (may be NULL) keyed by source, options and hiprtc version. */
int op2hip_compile(const char* src, const char* const* opts, int nopts, const char* cache_dir,
                   void** code_out, size_t* code_size, int* cache_hit);
void op2hip_free_host(void* p);
int op2hip_module_load(const void* code, size_t size, void** mod_out);
int op2hip_module_destroy(void* mod);
int op2hip_module_function(void* mod, const char* name, void** fn_out);
/* Device address and size of a module global (a kernel Const).
 * Replaces the per-call Const copy-in of pyop2/host.py:271-277. */
int op2hip_module_global(void* mod, const char* name, void** dptr, size_t* bytes);
/* Copy `bytes` host bytes to a module global on `stream`, then wait. */
int op2hip_set_global(void* dptr, const void* src, size_t bytes, void* stream);
/* Store up to 256 host bytes at `dptr` in `stream` order without waiting
 * (a kernel carrying them as its argument; capturable in a graph).  A
 * Global's value reaching the device: replaces the reduction-buffer
 * upload of pyop2/cuda.py:327-339 (gpuarray.to_gpu / fill), which the
 * reference follows with a stream synchronize per reduction
 * (pyop2/cuda.py:353-356). */
int op2hip_store_small(void* dptr, const void* src, size_t bytes, void* stream);
/* Compile (hiprtc, cached), load and resolve `entry` in one call; the
 * module belongs to the caller.  The JITModule compile+lookup of
 * pyop2/cuda.py:602-651 as one entry point. */
int op2hip_module_compile(const char* src, const char* entry, const char* const* opts, int nopts,
                          const char* cache_dir, void** mod_out, void** fn_out);
/* Set a module global (a Const, pyop2/base.py:857-926) by name. */
int op2hip_module_set_const(void* mod, const char* name, const void* host, size_t bytes,
                            void* stream);

/* Launch a generated loop kernel whose parameters are all 8 bytes wide:
 * argv[i] is parameter i (pointer or 64-bit integer).
 * Replaces the wrapper call fun(start, end, args...) of
 * pyop2/sequential.py:96-117 and the per-colour launches of
 * pyop2/cuda.py:757-800. */
int op2hip_launch(void* fn, unsigned grid, unsigned block, unsigned lds, void* stream,
                  const uint64_t* argv, int nargs);
int op2hip_function_attrs(void* fn, int* num_regs, int* local_bytes, int* max_threads);

/* Stream-ordered timing. */
int op2hip_event_create(void** ev);
int op2hip_event_destroy(void* ev);
int op2hip_event_record(void* ev, void* stream);
int op2hip_event_elapsed_ms(void* start, void* stop, float* ms);
int op2hip_stream_sync(void* stream);

/* ---- plan ----------------------------------------------------------------------
 * Replaces pyop2/plan.pyx:68-456 (Plan: partition, staging, two-level
 * colouring; properties plan.pyx:378-455).  Race maps are host int32 arrays
 * (row-major, dims[i] columns, at least nelems rows); idxs[i] the column
 * used (-1 all); spaces[i] the target index space of map i, of size
 * space_size[spaces[i]].  reorder = 1 orders the iteration set into compact
 * regions grown breadth-first through the race maps before partitioning
 * (two-level colouring: blocks, then threads inside a block, as plan.pyx);
 * reorder = 2 colours the elements themselves (greedy, in element order) and
 * cuts each colour's run into blocks: one thread colour per block, one
 * block colour per element colour. */
int op2hip_plan_build(int nelems, int part_size, int reorder, int nmaps, const int* const* maps,
                      const int* dims, const int* idxs, const int* spaces, int nspaces,
                      const int* space_size, void** plan_out);
/* field: 0 nelems, 1 nblocks, 2 ncolors, 3 max block size, 4 max thread colours */
int op2hip_plan_scalar(void* plan, int field, int64_t* out);
/* which: 0 elems, 1 offset, 2 nelems, 3 thrcol, 4 nthrcol, 5 blkcol, 6 blkmap,
 * 7 ncolblk (host int32 destination of the array's length) */
int op2hip_plan_array(void* plan, int which, int* dst);
int op2hip_plan_destroy(void* plan);
/* Conflict-freedom check (the invariant of test/unit/test_coloring.py:104-112). */
int op2hip_plan_verify(int nelems, int nblocks, const int* elems, const int* offset,
                       const int* nelem, const int* thrcol, const int* blkcol, int nmaps,
                       const int* const* maps, const int* dims, const int* idxs,
                       const int* spaces, int nspaces, const int* space_size, int* ok);

/* The coloured plan built on the device (the backend's default; the host
 * build above is kept as its alternative).  mode 0: two-level colouring of
 * blocks of part_size consecutive elements (block colours, then thread
 * colours inside each block greedily in element order, as plan.pyx); mode
 * 2: element colours, each colour's elements cut into blocks.  Colours by
 * Jones-Plassmann-Luby rounds with fixed hash priorities (deterministic).
 * maps: host array of device int32 map pointers; the other arguments as
 * op2hip_plan_build's.  work: op2hip_plan_workspace() bytes kept until
 * fill or discard; thrcol (device, nelems) receives mode 0's thread
 * colours.  dbuild returns nblocks and ncolors; dfill writes elems and
 * thrcol (device, nelems) and offset, nelem, nthrcol, blkcol, blkmap
 * (host, nblocks) and ncolblk (host, ncolors) with the meaning of
 * op2hip_plan_array's, and releases the handle.  Both wait for the stream. */
long long op2hip_plan_workspace(int nelems, int part_size, int nac, const int* space_size,
                                int nspaces);
int op2hip_plan_dbuild(int nelems, int part_size, int mode, int nmaps, const int* const* maps,
                       const int* dims, const int* idxs, const int* spaces, int nspaces,
                       const int* space_size, void* work, long long work_bytes, int* thrcol,
                       void** handle, int* nblocks_out, int* ncolors_out, void* stream);
int op2hip_plan_dfill(void* handle, int* elems, int* thrcol, int* offset, int* nelem,
                      int* nthrcol, int* blkcol, int* blkmap, int* ncolblk, void* stream);
int op2hip_plan_ddiscard(void* handle);

/* Owner-computes plan for loops whose conflicting writes are increments into
 * one target set (also replaces pyop2/plan.pyx:68-456 for those loops).
 * Targets are owned by blocks in contiguous ranges of k; a block runs every
 * element naming one of its targets (ascending order) and sums each owned
 * target's increments in that order.  Arrays (which): 0 blk_t0, 1 blk_t1,
 * 2 ex_off, 3 exec (bit 31: not primary), 4 tp, 5 contrib (slot*nac + ac),
 * 6 tg_off, 7 tgt (targets receiving increments, per block), 8 rng_lo,
 * 9 rng_n (a block's primary elements as one run when consecutive: slots
 * [0, rng_n) are elements rng_lo + slot, exec lists slots rng_n on).  Scalars
 * (field): 0 nactive, 1 max_exec, 2 nac, 3 exec length, 4 contrib length,
 * 5 tgt length.  elem_order (NULL: index order) is the order increments
 * are summed in. */
int op2hip_owner_plan_build(int nelems, int ntargets, int nmaps, const int* const* maps,
                            const int* dims, const int* idxs, int k, const int* elem_order,
                            void** plan_out);
int op2hip_owner_plan_scalar(void* plan, int field, int64_t* out);
/* The same owner plan built on the device by sort and scan (the form the
 * backend uses; the host build above is its restatement in tests).  Also
 * replaces pyop2/plan.pyx:68-456 for owner-computes loops.  maps: host array
 * of device int32 map pointers (rows in the loop's element order); rank
 * (device, nelems ints, may be NULL: the identity) orders each target's
 * increments (the host build's elem_order, inverted).  work: a device buffer
 * of op2hip_owner_plan_workspace() bytes, kept until fill or discard.  Phase
 * 1 (dbuild) writes tp (device, ntargets + 1) and waits for the stream once
 * to return scalars[6]: nactive, exec length, tgt length, max_exec, largest
 * contribution count of a block, sum over blocks of elements executed.  The
 * caller allocates the device arrays by these sizes; phase 2 (dfill) writes
 * blk_t0, blk_t1, rng_lo, rng_n (nactive), ex_off, tg_off (nactive + 1),
 * exec, tgt (tgt length + 1: one padding entry), tg_cb (each listed
 * target's first contribution, tgt length + 2) and contrib (nelems * race
 * columns entries, uint16 when c16 else int32), entry for entry the host
 * build's arrays. */
long long op2hip_owner_plan_workspace(int nelems, int ntargets, int nac, int k);
int op2hip_owner_plan_dbuild(int nelems, int ntargets, int nmaps, const int* const* maps,
                             const int* dims, const int* idxs, int k, const int* rank,
                             void* work, long long work_bytes, int* tp, void** handle,
                             long long* scalars, void* stream);
int op2hip_owner_plan_dfill(void* handle, const int* tp, int* blk_t0, int* blk_t1, int* rng_lo,
                            int* rng_n, int* ex_off, int* exec, int* tg_off, int* tgt,
                            int* tg_cb, void* contrib, int c16, void* stream);
int op2hip_owner_plan_ddiscard(void* handle);
int op2hip_owner_plan_array(void* plan, int which, int* dst);
int op2hip_owner_plan_destroy(void* plan);
/* Scatter program of an owner-computes Mat assembly (no counterpart in the
 * reference, whose addto_scalar/addto_vector, pyop2/mat_utils.cxx:4-28,
 * search each entry's row at add time): per node pair of the Mat's pattern,
 * the (block-local contribution, column position) codes c*e1+j that add
 * into it, in sequential order.  poff[npairs+1] and prog[ncontrib*e1]
 * (uint16 when wide == 0, else int32) are caller-allocated host arrays. */
int op2hip_mat_program(int nactive, const int* blk_t0, const int* blk_t1, const int* ex_off,
                       const int* exec, const int* rng_lo, const int* rng_n, const int* tp,
                       const int* contrib, int nac,
                       const int* cmap, int carity, int e1, int nnodes, const int* rowptr,
                       const int* colidx, int rd, int cd, int wide, int* poff, void* prog);

/* The same program built on the device from the device owner plan's
 * arrays (contrib uint16 when c16, else int32) and the device CSR: poff
 * (device, npairs + 1) and prog (device, nelems * nac * e1 codes, uint16
 * unless wide).  One thread per row node walks its contributions in plan
 * order, so each node pair receives its entries in sequential order
 * without atomics.  work: op2hip_mat_program_workspace(npairs) bytes.
 * Waits for the stream once (to report an element naming an entry outside
 * the pattern). */
long long op2hip_mat_program_workspace(long long npairs);
int op2hip_mat_program_device(int nactive, const int* blk_t0, const int* blk_t1,
                              const int* ex_off, const int* exec, const int* rng_lo,
                              const int* rng_n, const int* tp, const void* contrib, int c16,
                              int nac, const int* cmap, int carity, int e1, int nnodes,
                              const int* rowptr, const int* colidx, int rd, int cd,
                              long long npairs, int wide, void* work, long long work_bytes,
                              int* poff, void* prog, void* stream);

/* Row-block plan of the owner Mat "rows" template (replaces, for Mat
 * assembly, the plan of plan.pyx:68-456 and the per-entry search of
 * mat_utils.cxx:4-28 / petsc_base.py:92-98).  Block b owns row nodes
 * [b*k, min(nnodes, (b+1)*k)) and runs the elements naming them, in
 * ascending user order (rank; NULL: identity); one thread per owned node
 * adds its records' entries into the block's CSR slice in sequential order.
 * Outputs (device, caller-allocated): tp[nnodes+1] record offsets per node,
 * xoff[nblk+1] and exec (nelems*e0, an upper bound) the blocks' elements,
 * code32[nelems*e0] codes slot*e0 + i0, pos[nelems*e0*ceil(e1/4)] column
 * positions as bytes.  scalars (host, 5): largest code, largest exec list,
 * largest CSR slice (entries), exec entries in all, flags (1: an entry
 * outside the pattern, 2: a row of more than 256 node columns, 4: an
 * inconsistent plan).  work: op2hip_mat_rows_workspace() bytes.  Waits for
 * the stream once. */
long long op2hip_mat_rows_workspace(int nelems, int nnodes, int e0, int k);
int op2hip_mat_rows_build(int nelems, int nnodes, const int* rmap, int rarity, int e0,
                          const int* cmap, int carity, int e1, const int* rank,
                          const int* rowptr, const int* colidx, int rd, int cd, int k, void* work,
                          long long work_bytes, int* tp, int* xoff, int* exec, int* code32,
                          unsigned* pos, long long* scalars, void* stream);
/* code32[0, n) narrowed to uint16 into out (every code below 65536). */
int op2hip_mat_rows_pack16(const int* code32, long long n, unsigned short* out, void* stream);

/* Per-block staging lists of a plan: replaces the staging pass of
 * pyop2/plan.pyx:112-212 (per block and indirect map, numpy.unique of the
 * targets its elements name: ind_map, loc_map, ind_sizes, ind_offs).  Block
 * b's elements are elems[xoff[b] .. xoff[b+1]) (device); its slots are
 * xoff[b]*dim .. xoff[b+1]*dim of ind_map and loc_map (ind_offs = xoff*dim):
 * ind_map holds its sorted distinct targets (ind_sizes[b] of them) from the
 * start of that range, loc_map per slot the position of the slot's target
 * in that list.  maxkeys bounds (xoff[b+1]-xoff[b])*dim and may be at most
 * op2hip_stage_max_keys() (LDS per workgroup); *err (device) gets bit 1 for
 * a block beyond maxkeys or nslots.  Enqueues only. */
int op2hip_stage_max_keys(void);
int op2hip_stage_lists(int nblocks, const int* xoff, const int* elems, const int* map, int dim,
                       int maxkeys, long long nslots, int* ind_map, unsigned short* loc_map,
                       int* ind_sizes, int* err, void* stream);
/* The lists end to end: so[nblocks+1] the exclusive scan of ind_sizes,
 * sn[so[b] ..] block b's list.  Enqueues only. */
int op2hip_stage_compact(int nblocks, const int* xoff, int dim, const int* ind_sizes,
                         const int* ind_map, int* so, int* sn, void* stream);

/* ---- data layout ----------------------------------------------------------------
 * No counterpart in the reference, which stores every Set in user order.
 * Orderings that renumber a Set's device storage for locality (user-visible
 * numbering is unchanged).  perm[pos] = element stored at position pos.
 * bfs: breadth-first order of a map's targets (adjacent when one iteration
 * element names both); bisect: recursive bisection of that adjacency by
 * breadth-first cuts down to ranges of `leaf` targets (every aligned range
 * a compact patch); first_touch: targets in order of first use by the
 * source elements visited in src_order (NULL: index order). */
int op2hip_order_bfs(int ntarget, int nelem, const int* map, int dim, int* perm);
int op2hip_order_bisect(int ntarget, int nelem, const int* map, int dim, int leaf, int* perm);
int op2hip_order_first_touch(int ntarget, int nsrc, const int* map, int dim,
                             const int* src_order, int* perm);

/* ---- sparsity ------------------------------------------------------------------
 * Replaces build_sparsity_pattern_seq (pyop2/sparsity_utils.h:10-13,
 * sparsity_utils.cxx:5-51) as called by build_sparsity
 * (pyop2/op_lib_core.pyx:544-596).  Two phases so the caller allocates:
 * build writes rowptr (device, nrows*rmult+1) and *nnz_out; fill writes
 * colidx (device, nnz) and nnz_rows (device, nrows*rmult, may be NULL) and
 * releases the handle.  rowmaps/colmaps: host arrays of device int32 map
 * pointers (at most 8 map pairs); nelems[m]: iteration-set size of pair m.
 * work: a device buffer of op2hip_sparsity_workspace() bytes (about
 * 4 * (4 * nrows + I * (1 + max cdims)), I the sum of nelems[m] * rdims[m]),
 * kept untouched until the handle is filled or discarded; the build
 * allocates no device memory.
 * Build waits for the stream once (nnz), fill once (its check that every
 * row recounted to the build's count).  Bit-exact to the reference's
 * sorted-set construction. */
long long op2hip_sparsity_workspace(int nrows, int nmaps, const long long* nelems,
                                    const int* rdims, const int* cdims);
int op2hip_sparsity_build(int rmult, int cmult, int nrows, int nmaps, const int* const* rowmaps,
                          const int* rdims, const int* const* colmaps, const int* cdims,
                          const long long* nelems, void* work, long long work_bytes,
                          int* rowptr, void** handle, long long* nnz_out, void* stream);
int op2hip_sparsity_fill(void* handle, const int* rowptr, int* colidx, int* nnz_rows,
                         void* stream);
int op2hip_sparsity_discard(void* handle);

/* ---- matrices and solver ---------------------------------------------------------
 * Mat storage is FP64 CSR on the Sparsity (pyop2/petsc_base.py:80-120,
 * createAIJWithArrays); element insertion (addto_scalar/addto_vector,
 * pyop2/mat_utils.h:8-10) happens inside the generated loops. */

/* y = A x.  Replaces PETSc MatMult inside KSP. */
int op2hip_spmv(int n, long long nnz, const int* rowptr, const int* colidx, const double* vals,
                const double* x, double* y, void* stream);

/* Zero rows except the diagonal, set to diag.
 * Replaces Mat.zero_rows -> MatZeroRowsLocal (pyop2/petsc_base.py:131-135). */
int op2hip_csr_zero_rows(int nlist, const int* rows, const int* rowptr, const int* colidx,
                         double* vals, double diag, void* stream);

/* Batched Jacobi-preconditioned CG from x = 0, PETSc KSPCG semantics per
 * system.  Replaces Solver.solve (pyop2/petsc_base.py:182-220; parameters
 * pyop2/base.py:1591-1601); a single solve is a batch of one.  A system is
 * rows [row0, row0+n) of a CSR (global column numbers) with its vectors;
 * reason is PETSc's KSPConvergedReason (2 rtol, 3 atol, -3 its, -4 dtol,
 * -8 indefinite, -9 nan).  Reductions are per 256-row chunk of a system,
 * summed in chunk order, so a system's result does not depend on the other
 * systems of the batch. */
typedef struct op2hip_cg_system {
  const int* rowptr;
  const int* colidx;
  const double* vals;
  const double* b;
  double* x;
  double* inv_diag;
  double* r;
  double* zp; /* z and p interleaved: zp[2i] = z_i, zp[2i+1] = p_i; 16-byte aligned */
  double* w;
  int row0;
  int n;
  int chunk0;
  int nchunks;
  double rtol, atol, dtol;
  int maxit;
  int state; /* 0 empty, 1 new, 2 running, 3 done */
  int no_precond; /* 0: Jacobi (PCJACOBI), 1: none (PCNONE) */
  double beta, betaold, dpi, dpiold, dp, rnorm0, ttol, alpha;
  int its, reason;
} op2hip_cg_system;

int op2hip_cgb_create(int max_sys, long long max_chunks, int hist_len, void** out);
int op2hip_cgb_destroy(void* h);
int op2hip_cgb_load(void* h, int nsys, const op2hip_cg_system* systems, void* stream);
int op2hip_cgb_step(void* h, int iters, void* stream);
/* The same step replayed from a hipGraph captured (on a stream of the
 * batch's own) once per chunk list and iteration count: the batching
 * layer's captured CG steps.  Results equal op2hip_cgb_step's. */
int op2hip_cgb_step_graph(void* h, int iters, void* stream);
int op2hip_cgb_read(void* h, op2hip_cg_system* out, void* stream);
int op2hip_cgb_history(void* h, int k, double* out, void* stream);
int op2hip_cgb_sizeof_system(void);
int op2hip_cgb_chunk_rows(void);
/* One solve (replaces Solver.solve, pyop2/petsc_base.py:182-220, for one
 * CSR on the device): work holds 5*n device doubles, 16-byte aligned (D^-1,
 * r, z and p interleaved over 2*n, w); its, rnorm and reason
 * (KSPConvergedReason) are host outputs.  Waits for the stream. */
int op2hip_pcg(int n, const int* rowptr, const int* colidx, const double* vals, const double* b,
               double* x, double* work, double rtol, double atol, double dtol, int maxit,
               int* its, double* rnorm, int* reason, void* stream);

/* op2hip_pcg without device allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results (bit for bit) as op2hip_pcg; the solver's descriptor, chunk list
 * and partial sums live in a static per-GPU array of about 1.6 MB instead,
 * so n is at most op2hip_pcg_noalloc_max_rows() (2^24) and larger systems
 * are refused with an error.  Calls that overlap on one GPU race on that
 * array.  Waits for the stream. */
int op2hip_pcg_noalloc(int n, const int* rowptr, const int* colidx, const double* vals,
                       const double* b, double* x, double* work, double rtol, double atol,
                       double dtol, int maxit, int* its, double* rnorm, int* reason,
                       void* stream);
int op2hip_pcg_noalloc_max_rows(void);

/* GMRES and BiCGStab for one CSR on the device (Solver.solve with
 * linear_solver 'gmres' or 'bcgs'/'bicgstab', pyop2/petsc_base.py:172-220;
 * the cusp backend's methods, pyop2/cuda.py:504-508): PETSc 3.3 KSPGMRES
 * (restarted every `restart` steps, classical Gram-Schmidt) or KSPBCGS,
 * with PCJACOBI (jacobi 1) or PCNONE (jacobi 0) on the left and the
 * preconditioned residual norm.  method 1 is GMRES, 2 BiCGStab.  op2hip_krylov_work gives the device workspace in
 * doubles (-1 for bad arguments); x is overwritten (zero initial guess).
 * reshist (device, hist_len doubles, may be NULL with hist_len 0) receives
 * the monitor's residual norm per iteration.  its, rnorm and reason
 * (KSPConvergedReason) are host outputs.  Waits for the stream.  The
 * iterations run as captured hipGraphs, kept for the last 8 distinct
 * argument lists and replayed by calls with the same arguments. */
long long op2hip_krylov_work(int method, int n, int restart);
int op2hip_krylov_solve(int method, int n, const int* rowptr, const int* colidx,
                        const double* vals, const double* b, double* x, double* work,
                        long long work_len, int restart, int jacobi, double rtol, double atol,
                        double dtol, int maxit, double* reshist, int hist_len, int* its,
                        double* rnorm, int* reason, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* OP2HIP_H */
