/* sawfem: MI355X (gfx950) HIP backend for the SawSim piezoelectric FEM frequency sweep.
 *
 * Plain C ABI: device pointers are passed as void*, sizes as int/int64_t, the stream as
 * a hipStream_t cast to void*.  Complex values are interleaved (re, im) doubles
 * ("complex128") unless stated.  Status codes (int return):
 *   0 success
 *   1 invalid argument
 *   2 plan limit exceeded (front count or order beyond what the plan encodes)
 *   3 HIP runtime error
 *   4 out of device memory (an allocation failed, or a caller workspace is too small)
 *   5 non-finite result (sawfem_sweep_chunk: a charge is NaN or infinite)
 *   6 success, with the complex128 refactorization used for some frequencies
 *     (sawfem_sweep_chunk; which ones is in its flags output)
 * sawfem_last_error() returns the thread-local message of the last failure.
 *
 * Reference interfaces replaced (SawSim, Python):
 *   saw2d/piezo_fem.py:283-367 assemble / :373-405 assemble_ME0  -> sawfem_elem_q9 + sawfem_gather_kmd
 *   saw2d/saw3d/assembly.py:230-327 form_stiffness_mass_3d         -> sawfem_elem_h27 + sawfem_gather_h27
 *   saw2d/saw3d/bc_solve.py:121-175 FrequencySolver3D (two-stage fold, active set, Q rows)
 *                                                                   -> sawfem_gather3 (fold as a representative map)
 *   saw2d/sweep.py:283-323 FrequencySolver.__init__/set_damping     -> sawfem_gather3 (Tᵀ K T, active extraction,
 *                                                                     Km@s, Kᵀs)
 *   saw2d/sweep.py:326-333, 349-353 _backend / PARDISO / SuperLU   -> sawfem_mf_create, sawfem_mf_factor,
 *                                                                     sawfem_mf_solve
 *   saw2d/sweep.py:337-342 rhs = -(Ks - w^2 Ms)[active] V          -> sawfem_rhs
 *   saw2d/sweep.py:362 Q = kts . disp                               -> sawfem_charge
 *   (no counterpart; refinement against the FP64 operator)          -> sawfem_residual, sawfem_axpy
 *   saw2d/sweep.py:335-364 FrequencySolver.solve (whole chunk)     -> sawfem_sweep_chunk
 */
#ifndef SAWFEM_H
#define SAWFEM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* sawfem_last_error(void);

/* K1: Q9 element matrices.  conn (ne*9 int32, QUADS9 order), xy (N*2 double, metres),
 * mat_idx/elem_tag (ne int32), per material C9 (nmat*81), E9 (nmat*27), EPS (nmat*9)
 * complex128 (the 6x6 / 3x6 tensors lifted to 9-component strain), rho (nmat double);
 * npml <= 8 PML regions given by host arrays tag, kind (bit0 x, bit1 y) and
 * par (6 doubles: dmax, n, xa, xp, ya, yp).  ndu = 2 (plane strain [ux|uy|phi]) or 3.
 * Writes Ke[e][A][B][a][b] (ne*(ndu+1)^2*81) and Me[e][a][b] (ne*81), complex128.
 * Replaces piezo_fem.py:248-350 (_element_geometry, element_alpha, block contractions). */
int sawfem_elem_q9(void* stream, int ne, int ndu, const int* conn, const double* xy, const int* mat_idx,
                   const int* elem_tag, const void* C9, const void* E9, const void* EPS, const double* rho,
                   int npml, const int* pml_tag, const int* pml_kind, const double* pml_par,
                   double omega_top, double scale, void* Ke, void* Me);

/* K2: global K, M (and D = K_uu of elements with dmask[e] != 0, if D != NULL) on the
 * structural CSR pattern, gathered per node pair in a fixed order (no atomics).
 * Replaces the COO -> CSR scatter of piezo_fem.py:352-365. */
int sawfem_gather_kmd(void* stream, int nnzn, int nd, int ndu, const int* pair_row, const int64_t* node_ptr,
                      const int64_t* gptr, const int64_t* gidx, const unsigned char* dmask, const void* Ke,
                      const void* Me, void* K, void* M, void* D);

/* K1' (Hex27): element matrices of the 27-node slab.  conn (ne*27, reference H27 node order),
 * xyz (N*3, metres), mat_idx/pml_idx per element (pml_idx -1: no stretch); record tensors
 * C (nrec*36), E (nrec*18), EPS (nrec*9) complex128 row-major, rho (nrec) real; npml <= 8 PML
 * parameter sets: kind (bit 0: x stretch, bit 1: z stretch) and par (6 doubles: dmax, n, xa, xp,
 * za, zp); the stretch is taken at local node 26 and frozen at omega_top.  scale = KM_SCALE.
 * Writes Ke[e][A*4+B][a*27+b] (ne*16*729) and Me[e][a*27+b] (ne*729), complex128.
 * Replaces saw3d/assembly.py:31-308 (_geometry, _b_matrix with the gamma_yz quirk, _bfi_matrix,
 * _stretch_alphas, _c/_e/_eps_stretch, _pml_fields, B^T D B products, conj(K_uphi)^T). */
int sawfem_elem_h27(void* stream, int ne, const int* conn, const double* xyz, const int* mat_idx, const int* pml_idx,
                    const void* C, const void* E, const void* EPS, const double* rho, int npml, const int* pml_kind,
                    const double* pml_par, double omega_top, double scale, void* Ke, void* Me);

/* K2' (Hex27): K, M (and D = K_uu of elements with dmask[e] != 0, if D != NULL) on the 27-node
 * structural pattern, 4 components, block-major slots; codes e*729 + i*27 + j per node pair.
 * Replaces the sparse products and _scatter_mass of saw3d/assembly.py:288-327. */
int sawfem_gather_h27(void* stream, int nnzn, const int* pair_row, const int64_t* node_ptr, const int64_t* gptr,
                      const int64_t* gidx, const unsigned char* dmask, const void* Ke, const void* Me, void* K, void* M,
                      void* D);

/* K12' (Hex27 backward): per element of elems (ne0) and frequency, the sensitivity sums of
 * l = s - T P_a^T lambda and the full vector d, contracted on the host with the tensor
 * derivatives: W[b] = (Hc 6x6, He 3x6, Hp 3x3, Hm), 64 complex128, already multiplied by the
 * element's stretch codes.  xt, lt: x and lambda as [n][Bp] (Bp >= B); pos_rep: active position
 * of each dof's representative (-1 if none); zmask: signal-electrode potentials; work:
 * B*ne0*64 complex128.  No counterpart in the reference (it has no Hex27 gradient); the
 * element operators are those of sawfem_elem_h27. */
int sawfem_adjoint_h27(void* stream, int ne0, const int* elems, const int* conn, const double* xyz, const int* pml_idx,
                       int npml, const int* pml_kind, const double* pml_par, double omega_top, int N,
                       const int* pos_rep, const unsigned char* zmask, double V, double scale, int B, int Bp,
                       const void* xt, const void* lt, void* work, void* W);

/* K3: d_i[j] = sum_{c in ptr[j]..ptr[j+1]} s_i[idx[c]] for up to three complex128 sources
 * sharing one list (s1/d1, s2/d2 may be NULL together).  Used for Tᵀ A T restricted to the
 * active set, the Km@s column and the charge weights (sweep.py:293-309). */
int sawfem_gather3(void* stream, int64_t nout, const int64_t* ptr, const int64_t* idx, const void* s0,
                   const void* s1, const void* s2, void* d0, void* d1, void* d2);

/* Host symbolic analysis (no GPU call): nested-dissection ordering of the active system and the
 * multifrontal plan arrays.  n active unknowns with pattern aa_indptr (n+1) / aa_indices (CSR,
 * int64); sv_of[i] in [0, nxy) groups unknowns into supervariables (merged mesh nodes, or slab
 * columns) with coordinates xy (nxy*2 doubles, x then y; the cuts are made on these), comp_of[i]
 * orders the unknowns within one; period > 0 makes x periodic with that period (0: not
 * periodic); leaf_weight is the largest sub-domain (in unknowns) that is not dissected further.
 * The result is read with sawfem_host_size / sawfem_host_copy (element counts; int32 or int64
 * arrays as listed) and freed with sawfem_host_destroy; sawfem_mf_create_sym builds the device
 * plan from it.  Replaces the per-frequency analysis phase of PARDISO / SuperLU
 * (sweep.py:326-333, 349-353), done once per cell here. */
enum {
  SAWFEM_SYM_M = 0,          /* int32 [nf]     front order */
  SAWFEM_SYM_P = 1,          /* int32 [nf]     pivots of the front */
  SAWFEM_SYM_ROWS_PTR = 2,   /* int64 [nf+1] */
  SAWFEM_SYM_ROWS = 3,       /* int32          front rows (active ids), pivots first, elimination order */
  SAWFEM_SYM_CHILD_PTR = 4,  /* int32 [nf+1] */
  SAWFEM_SYM_CHILD_IDS = 5,  /* int32 */
  SAWFEM_SYM_PARENT = 6,     /* int32 [nf]     -1 for a root */
  SAWFEM_SYM_LEVEL = 7,      /* int32 [nf]     0 for a leaf */
  SAWFEM_SYM_INV_PTR = 8,    /* int64          one map per (front, child) pair */
  SAWFEM_SYM_INV = 9,        /* int32          parent row -> child contribution index or -1 */
  SAWFEM_SYM_LVL_PTR = 10,   /* int32 [nlev+1] */
  SAWFEM_SYM_LVL_FRONTS = 11,/* int32 [nf]     fronts by level, then id */
  SAWFEM_SYM_ENT_FRONT = 12, /* int32 [nnz]    owner front of each A_aa entry, sorted */
  SAWFEM_SYM_ENT_SLOT = 13,  /* int64 [nnz]    the entry's CSR position */
  SAWFEM_SYM_ENT_LOC = 14,   /* int32 [nnz]    column-major offset in the owner front */
  SAWFEM_SYM_ENT_LVL_PTR = 15,/* int64 [nlev+1] */
  SAWFEM_SYM_PERM = 16,      /* int64 [n]      active unknowns in elimination order */
  SAWFEM_SYM_PIV_SV_PTR = 17,/* int64 [nf+1] */
  SAWFEM_SYM_PIV_SV = 18,    /* int64 [nsv]    supervariables in elimination order */
  SAWFEM_SYM_ROWS_SV_PTR = 19,/* int64 [nf+1] */
  SAWFEM_SYM_ROWS_SV = 20,   /* int64          front rows as supervariables */
  SAWFEM_SYM_SV_IDS = 21     /* int64 [nsv]    sv_of value of each supervariable (ascending) */
};
int sawfem_analyse(void** sym, int64_t n, const int64_t* aa_indptr, const int64_t* aa_indices, const int64_t* sv_of,
                   const int64_t* comp_of, int64_t nxy, const double* xy, double period, int64_t leaf_weight);
int sawfem_mf_create_sym(void** plan, const void* sym);

/* Host Bloch reduction layout (no GPU call): the structural pattern of
 * A_aa = (T^T K T)[active, active] and the gather lists the device gathers use
 * (sawfem_gather3), over the K slots of the structural pattern.  K's slots: node adjacency
 * node_ptr (N+1) / node_idx, nd components per node, slot = a*nd*nnz_node + nd*node_ptr[i] +
 * b*deg[i] + (k - node_ptr[i]) for component block (a, b) of node pair k of row node i; dof of
 * (a, node) = a*N + node, or (interleaved != 0, nd = 4) 3*node + a for a < 3 and 3N + node.
 * rep[d]: Bloch representative of dof d (rep[rep[d]] = rep[d]); pos[d]: active position of d or
 * -1; is_zheng / is_fu: 0/1 flags of the driven-electrode and grounded dofs.  Output arrays
 * (int64, read with sawfem_host_size / sawfem_host_copy, freed with sawfem_host_destroy):
 *   AA_INDPTR / AA_INDICES  the A_aa pattern (CSR, columns ascending);
 *   AA_PTR / AA_IDX         per A_aa entry, the K slots whose merged row and column map to it;
 *   RHS_PTR / RHS_IDX       per active row, the K slots with merged row there and merged column
 *                           on the electrode (the Km@s column, sweep.py:303);
 *   KTS_PTR / KTS_IDX       charge weights on the unmerged K (sweep.py:305): per active position,
 *                           slots of an electrode row whose merged column is that position (column
 *                           neither electrode nor grounded); entry n_active: slots of an electrode
 *                           row and electrode column.
 * Every list is in ascending slot order.  Replaces the per-cell set-up of
 * FrequencySolver.__init__ (sweep.py:283-314: union of slave into master rows and columns,
 * active extraction). */
enum {
  SAWFEM_BLOCH_AA_INDPTR = 32,
  SAWFEM_BLOCH_AA_INDICES = 33,
  SAWFEM_BLOCH_AA_PTR = 34,
  SAWFEM_BLOCH_AA_IDX = 35,
  SAWFEM_BLOCH_RHS_PTR = 36,
  SAWFEM_BLOCH_RHS_IDX = 37,
  SAWFEM_BLOCH_KTS_PTR = 38,
  SAWFEM_BLOCH_KTS_IDX = 39
};
int sawfem_bloch_layout(void** out, int nd, int64_t N, int interleaved, const int64_t* node_ptr,
                        const int64_t* node_idx, const int64_t* rep, const int64_t* pos, int64_t n_active,
                        const unsigned char* is_zheng, const unsigned char* is_fu);

/* Host (no GPU call): tslot[e] = position of the transposed entry (c, r) of every entry e =
 * (r, c) of a structurally symmetric CSR pattern with ascending columns in every row (status 1
 * otherwise).  Used by the transposed residual and the paired adjoint; no counterpart in the
 * reference (it solves the adjoint nowhere). */
int sawfem_transpose_slots(int64_t n, const int64_t* indptr, const int64_t* indices, int64_t* tslot);

/* Arrays of a host result (sawfem_analyse, sawfem_bloch_layout): element count of array
 * `what` (-1 for an id the result does not have), copy into dst, free the result. */
int64_t sawfem_host_size(const void* res, int what);
int sawfem_host_copy(const void* res, int what, void* dst);
int sawfem_host_destroy(void* res);

/* Multifrontal plan from the host symbolic analysis (all pointers are host memory). */
int sawfem_mf_create(void** plan, int n, int nf, const int* m, const int* p, const int64_t* rows_ptr,
                     const int* rows, const int* child_ptr, const int* child_ids, const int64_t* inv_ptr,
                     const int* inv, int nlev, const int* lvl_ptr, const int* lvl_fronts, int64_t nent,
                     const int* ent_front, const int64_t* ent_slot, const int* ent_loc, const int64_t* ent_lvl_ptr);
int sawfem_mf_destroy(void* plan);
int64_t sawfem_mf_arena_elems(void* plan);   /* complex entries of one frequency's fronts */
int64_t sawfem_mf_uarena_elems(void* plan);  /* complex128 entries of one frequency's update vectors */
int sawfem_mf_perturbed(void* plan, int reset);
/* Pivot guard of sawfem_mf_factor (static pivoting with perturbation, no row exchanges): a
 * pivot of row g with |pivot| < sqrt(u_T) * s_g(w) is replaced by sqrt(u_T) * s_g(w) with its own
 * phase (u_T the unit roundoff of the front precision, s_g(w) = max|K_g.| + w^2 max|M_g.| +
 * w |beta| max|D_g.| the row scale of A(w)).  Reports the pivots replaced and the smallest
 * |pivot| / s_g since the last reset (reset != 0 clears both).  Replaces the perturbation
 * report of the PARDISO backend (sweep.py:349-353), which SawSim does not read. */
int sawfem_mf_pivot_stats(void* plan, int reset, int64_t* perturbed, double* min_ratio);

/* Fronts of order <= max_m are factored by one fused kernel each (assembly, partial LU and
 * diagonal-block inverses with the front in LDS) and solved by one workgroup each; the others
 * by the panel kernels.  Per precision (prec 0: complex64, 1: complex128); max_m <= 0 (the
 * default) turns the fused path off; requests are clipped to what the device LDS holds (at
 * most 128).  Returns the threshold in use, or -1.  No counterpart in the reference (a
 * kernel-selection knob of this backend). */
int sawfem_mf_set_small(void* plan, int prec, int max_m);
/* In the forward passes, fronts of order <= max_m are solved by one workgroup each (pivot
 * triangle and off-diagonal panel in one kernel) instead of the pivot-triangle + panel kernel
 * pair; 0 keeps the pair.
 * Returns the order in use (capped at 1024), or -1.  A tuning knob of the frontal solves
 * (part of the PARDISO/SuperLU replacement, sweep.py:349-353); no separate reference interface. */
int sawfem_mf_set_solve_full(void* plan, int max_m);
/* In the factorization, the 32-wide pivot panels of front-size buckets whose largest front has
 * order <= max_m are factored by one kernel per panel (diagonal block LU and inverses and both
 * panel solves, one wave per front and frequency) instead of the diagonal-block kernel and two
 * MFMA panel-solve GEMMs; 0 (the default) keeps the latter.  Returns the order in use, or -1.
 * A kernel-selection knob of the PARDISO/SuperLU replacement (sweep.py:349-353). */
int sawfem_mf_set_panel_lu(void* plan, int max_m);
/* In the solves (not transposed), fronts of the one-workgroup set (sawfem_mf_set_solve_full) whose
 * size bucket's largest factor block (m * p entries of the front precision) fits max_bytes are
 * solved with that block staged in LDS by one coalesced load, forward and backward; 0 (the
 * default) keeps the pivot-triangle kernels.  Clipped to the device's LDS per workgroup; returns
 * the bytes in use, or -1.  A kernel-selection knob of the PARDISO/SuperLU replacement. */
int sawfem_mf_set_tri_lds(void* plan, int max_bytes);

/* Factor A(w_b) = K + i w_b beta D - w_b^2 M for B frequencies (omega: device doubles)
 * into arena (arena_elems * B entries; complex64 if prec == 0, complex128 if prec == 1).
 * Kaa/Maa/Daa: complex128 values on the active pattern (Daa may be NULL).  nb: panel width.
 * Work is ordered on `stream`: the independent size buckets of a tree level also run on two
 * side streams of the plan (default priority, created on first use), forked from and joined
 * back to `stream` within the call, so a caller sees ordinary stream semantics.  The same holds
 * for sawfem_mf_solve.  A plan is used by one host thread at a time. */
int sawfem_mf_factor(void* plan, void* stream, int prec, int B, const double* omega, double beta, const void* Kaa,
                     const void* Maa, const void* Daa, void* arena, int nb);

/* sawfem_mf_factor as one HIP graph launch: the factorization's launch sequence (every tree
 * level, the side-stream forks and joins) is captured once per argument set (prec, B, the
 * omega/Kaa/Maa/Daa/arena pointers, beta, nb) and replayed with hipGraphLaunch on `stream`;
 * the values behind the pointers are read when the graph runs, so a caller that keeps its
 * buffers (new frequencies written into the same omega buffer, the same workspace) replays one
 * graph per chunk size.  Up to four graphs per plan (least recently used replaced); changing a
 * kernel-selection knob drops them.  With per-launch profiling on it runs eagerly.  Same results
 * as sawfem_mf_factor bit for bit.  The north star's graph-captured batched factorization; in the
 * reference the per-frequency backend call (SawSim saw2d/sweep.py:349-353). */
int sawfem_mf_factor_graph(void* plan, void* stream, int prec, int B, const double* omega, double beta,
                           const void* Kaa, const void* Maa, const void* Daa, void* arena, int nb);
/* Nodes (kernel launches and stream dependencies) of the graph the last sawfem_mf_factor_graph
 * replayed, or -1.  No counterpart in the reference. */
int64_t sawfem_mf_graph_nodes(void* plan);

/* Forward/backward substitution with the factors of sawfem_mf_factor; x: B x nrhs x n
 * complex128 (nrhs = 1 or 2 right-hand sides per frequency, sharing its factors: each factor
 * entry is read once for all of them), right-hand sides in, solutions out.
 * uarena: uarena_elems * B * nrhs complex128 scratch.
 * transpose != 0 solves A^T x = b with the same factors (U^T then L^T; adjoint solve).
 * nb: the panel width the arena was factored with (the diagonal blocks store the inverses of
 * their triangular factors, so the solve has to walk the same blocks).  bsel: device list of
 * nsel frequencies of the chunk to solve (refinement steps skip converged frequencies), or
 * NULL for all B. */
int sawfem_mf_solve(void* plan, void* stream, int prec, int B, int nrhs, const void* arena, void* uarena, void* x,
                    int transpose, int nb, const int* bsel, int nsel);
/* sawfem_mf_solve with factors shared between frequencies: right-hand side b (of B) is solved
 * with the factors in slot fslot[b] (device array of B ints in [0, Bf)) of an arena factored for
 * Bf frequencies.  Refined against each frequency's own FP64 operator, the factors of a nearby
 * frequency act as a preconditioner (a few more correction steps instead of a factorization).
 * No counterpart in the reference (it factors every frequency). */
int sawfem_mf_solve_shared(void* plan, void* stream, int prec, int B, int nrhs, const void* arena, int Bf,
                           const int* fslot, void* uarena, void* x, int transpose, int nb, const int* bsel, int nsel);

/* sawfem_mf_solve_shared for uniform groups, complex64 factors: frequencies b of the chunk
 * ([B][nrhs][n] vectors in x) share the factors of arena slot b / G (the arena holds ceil(B/G)
 * factorizations).  One workgroup solves the vectors of several frequencies of a group
 * together (up to the block size set by sawfem_mf_set_group_rc, default 16 vectors), so each
 * factor entry is read once for all of them.  gsel (device int[ngsel], or NULL for all):
 * the full groups (G frequencies) to solve; tail (0/1): whether the last, partial group
 * (frequencies G floor(B/G) .. B-1) is solved.  The vectors of unselected groups are left
 * untouched; every frequency of a selected group is solved (a zero right-hand side gives a zero
 * solution).  No counterpart in the reference (it factors every frequency). */
int sawfem_mf_solve_groups(void* plan, void* stream, int B, int nrhs, const void* arena, int G, void* uarena,
                           void* x, int transpose, int nb, const int* gsel, int ngsel, int tail);
/* sawfem_mf_solve_groups (forward + backward) on vectors stored row-major: entry i of vector
 * b nrhs + r at x[i ldx + b nrhs + r], ldx >= B nrhs, so the vectors of one row are contiguous
 * and the solve's gathers and scatters read whole lines; the update vectors in uarena are stored
 * row-major as well.  The same operations as sawfem_mf_solve_groups, so the same results bit for
 * bit.  No counterpart in the reference. */
int sawfem_mf_solve_groups_rows(void* plan, void* stream, int B, int nrhs, const void* arena, int G, void* uarena,
                                void* x, int64_t ldx, int nb, const int* gsel, int ngsel, int tail);
/* Largest block of vectors one workgroup of sawfem_mf_solve_groups takes (1..16; the block
 * size used is the largest R 2^k within it that divides G nrhs and fits the LDS).  Returns the
 * value in use or -1.  No counterpart in the reference. */
int sawfem_mf_set_group_rc(void* plan, int rc_max);
/* sawfem_mf_solve_groups kernel selection: the one-workgroup solve buckets use the LDS-staged
 * block kernel where the padded vectors of blocks of 4 vectors fit lds_bytes of LDS (default
 * 48 KiB), the other buckets the block kernels for the pivot rows and the off-diagonal panels;
 * lds_bytes = 0 turns both off.  The pivot-triangle and panel kernels that remain (blocks that do
 * not fit, a partial last group, transposed solves) take blocks of at most rc_max vectors
 * (default 4).  Returns rc_max in use or -1.  No counterpart in the reference. */
int sawfem_mf_set_group_tri(void* plan, int rc_max, int lds_bytes);
/* sawfem_mf_solve_groups: the block kernels for pivot rows and off-diagonal panels take the
 * two-kernel buckets whose largest pivot block is at most max_p (default 1024; 0: none); larger
 * ones keep the pivot-triangle + panel pair.  Returns the value in use or -1.  No counterpart in
 * the reference. */
int sawfem_mf_set_group_big(void* plan, int max_p);

/* Per-kernel HIP-event timing of the factor/solve launches (no counterpart in the reference;
 * used by the benchmark): enable 1 times every launch (factorizations then run launch by launch,
 * not as graphs), 2 only the grouped solve kernel (sawfem_mf_profile_grp; graphs allowed), 0
 * turns it off.  Kinds: 0 init, 1 entries, 2 diag, 3 trsm, 4 gemm-L, 5 gemm-S, 6 fwd, 7 bwd,
 * 8 fused small fronts.  profile_read waits for the recorded events and returns totals in
 * ms[9] / count[9]. */
int sawfem_mf_profile(void* plan, int enable);
int sawfem_mf_profile_read(void* plan, double* ms, int64_t* count);
/* The same totals per level of the assembly tree (leaves first), ms[level * 9 + kind], from the
 * last profile_read; returns the number of levels written (at most max_levels) or -1. */
int sawfem_mf_profile_levels(void* plan, double* ms, int max_levels);
/* The grouped block-solve kernel (k_grp_lds on whole fronts, every block size and kind) since
 * profiling was turned on: summed HIP-event time of its launches (ms), their algorithmic bytes
 * (factor entries read once per block, vectors gathered and written) and the launch count.  Read
 * it before sawfem_mf_profile_read, which releases the events.  No counterpart in the reference. */
int sawfem_mf_profile_grp(void* plan, double* ms, double* bytes, int64_t* launches);

/* out[b][i] = -V (rK[i] + i w_b beta rD[i] - w_b^2 rM[i])  (sweep.py:338-342). */
int sawfem_rhs(void* stream, int n, int B, const double* omega, double beta, double V, const void* rK,
               const void* rM, const void* rD, void* out);

/* r[b] = rhs[b] - A(w_b) x[b] in FP64 on the active CSR pattern (indptr int64, indices int32);
 * with tslot != NULL (slot of the transposed entry, int64) the operator is A^T. */
int sawfem_residual(void* stream, int n, int B, const int64_t* indptr, const int* indices, const int64_t* tslot,
                    const void* Kaa, const void* Maa, const void* Daa, const double* omega, double beta,
                    const void* x, const void* rhs, void* r);

/* Same residual in the frequency-contiguous layout: xt, rhst, rt are n x Bp complex128
 * (row-major, Bp >= B a multiple of 64); one wave per row and 64 frequencies. */
int sawfem_residual_t(void* stream, int n, int B, int Bp, const int64_t* indptr, const int* indices,
                      const int64_t* tslot, const void* Kaa, const void* Maa, const void* Daa, const double* omega,
                      double beta, const void* xt, const void* rhst, void* rt);

/* sawfem_residual_t over ng row groups (gptr int32[ng + 1], grows int32: the rows of group g
 * are grows[gptr[g] .. gptr[g + 1]), at most 4, with identical column lists, e.g. the active
 * unknowns of one merged mesh node; every row in exactly one group).  One wave per group and 64
 * columns reads each x row once for the group's rows; per row the result equals
 * sawfem_residual_t's bit for bit.  B columns c of the [n][Bp] arrays xt, rhst, rt, stored at
 * column cmap[c] (device int, or NULL: column c, then Bp >= B); omega is indexed by the stored
 * column; the other columns are not touched.  rhst may be NULL (then rt = -A x).  Replaces the
 * same residual as sawfem_residual (the FP64 A(w) x of SawSim saw2d/sweep.py:335-364's solve,
 * here the refinement residual). */
int sawfem_residual_sv(void* stream, int ng, int B, int Bp, const int* cmap, const int* gptr, const int* grows,
                       const int64_t* indptr, const int* indices, const int64_t* tslot, const void* Kaa,
                       const void* Maa, const void* Daa, const double* omega, double beta, const void* xt,
                       const void* rhst, void* rt);

/* out (C rows, leading dimension ld_out) = transpose of in (R rows, leading dimension ld_in). */
int sawfem_transpose(void* stream, int R, int C, int ld_in, int ld_out, const void* in, void* out);
/* Mapped transpose: for r < R, c < C, out[omap[c] * ld_out + r] = add[omap[c] * ld_out + r] +
 * in[imap[r] * ld_in + c] (imap, omap: device int row maps, NULL for the identity; add: NULL for
 * none).  Gathers the active vectors of a refinement step into the frequency-major layout of
 * sawfem_residual_sv and scatters rhs - A x back to their rows.  No counterpart in the reference. */
int sawfem_transpose_map(void* stream, int R, int C, int ld_in, int ld_out, const void* in, void* out, const int* imap,
                         const int* omap, const void* add);

/* Q[b] = sum_i kts_a[i] x[b][i] + V kts_z  (sweep.py:356-362). */
int sawfem_charge(void* stream, int n, int B, const void* kts_a, double kz_re, double kz_im, double V,
                  const void* x, void* Q);

/* out[k] = sum_i a[b sa + i] w[i] c[b sc + i] for k < nb, b = sel[k] (device int, or k when sel is
 * NULL): complex128 a, c (strides sa, sc in complex entries), real weights w (NULL: 1), no
 * conjugation.  The refinement's duality products lambda^T D b and dual-weighted residuals
 * lambda^T D r in one pass over the two vectors; the quantities of SawSim's charge reduction
 * (saw2d/sweep.py:360-364, Q = kts . disp) carried over to the adjoint. */
int sawfem_wdot(void* stream, int n, int nb, const void* a, int64_t sa, const double* w, const void* c, int64_t sc,
                const int* sel, void* out);

/* sawfem_charge and sawfem_wdot on vectors stored row-major (entry i of column j at [i ld + j]):
 * Q[k] = sum_i kts_a[i] x[i][cols[k]] + V kts_z, out[k] = sum_i a[i][acol[k]] w[i] c[i][ccol[k]]
 * for k < nc (column maps: device int, NULL for k).  Summed in the order of sawfem_charge /
 * sawfem_wdot, so equal to them bit for bit on the same vectors.  work: 1024 nc complex128
 * partial sums.  No counterpart in the reference (sweep.py:356-364's charge reduction). */
int sawfem_charge_rows(void* stream, int n, int nc, int64_t ld, const void* kts_a, double kz_re, double kz_im,
                       double V, const void* x, const int* cols, void* work, void* Q);
int sawfem_wdot_rows(void* stream, int n, int nc, int64_t ld, const void* a, const int* acol, const double* w,
                     const void* c, const int* ccol, void* work, void* out);
/* Both duality products of the paired refinement in one pass, on row-major arrays whose columns
 * pair up per frequency (2b forward, 2b + 1 adjoint): out1[b] = sum_i X[i][2b + 1] w[i] C[i][2b],
 * out2[b] = sum_i X[i][2b] w[i] C[i][2b + 1] for b < nb (ld >= 2 nb), each equal bit for bit to
 * sawfem_wdot on the same vectors.  work: 2048 nb complex128.  No counterpart in the reference. */
int sawfem_wdot_pairs_rows(void* stream, int n, int nb, int64_t ld, const void* X, const double* w, const void* C,
                           void* work, void* out1, void* out2);
/* One update of the paired refinement on row-major column pairs, with its measurements, in one
 * pass: X[i][c] += R[i][c] for c < 2 nb, then Q[c] = sum_i kts_a[i] X[i][c] + V kts_z (c < 2 nb)
 * and mu[b] = sum_i X[i][2b + 1] w[i] rhs[i][2b] (b < nb), each equal bit for bit to sawfem_axpy,
 * sawfem_charge and sawfem_wdot run one after another.  work: 3072 nb complex128.  No counterpart
 * in the reference. */
int sawfem_rows_update_measure(void* stream, int n, int nb, int64_t ld, void* X, const void* R, const void* kts_a,
                               double kz_re, double kz_im, double V, const double* w, const void* rhs, void* work,
                               void* Q, void* mu);

/* x += d (n complex128). */
int sawfem_axpy(void* stream, int64_t n, void* x, const void* d);

/* K12 backward element kernel for dY11/dp: per frequency, the PML-weighted sensitivity
 * tensors of l = s - T P_a^T lambda and the displacement d over the elements listed in
 * elems (the parameterised material): W[b] = [Hc (NR x NR), He (2 x NR), Hp (2 x 2)],
 * NR = 2 ndu.  pos_rep: active index of each dof's merge representative or -1; zmask:
 * signal-electrode dofs (value V in d, 1 in s).  work: B*ne0*NT complex128 scratch.
 * Replaces the reference's central finite differences of the sweep (2P+1 solves per point). */
int sawfem_adjoint_elem(void* stream, int ne0, const int* elems, const int* conn, const double* xy,
                        const int* elem_tag, int npml, const int* pml_tag, const int* pml_kind, const double* pml_par,
                        double omega_top, int ndu, int N, const int* pos_rep, const unsigned char* zmask, double V,
                        double scale, int B, int n, const void* x, const void* lam, void* work, void* W);

/* Active operator of one cell for sawfem_sweep_chunk (all pointers device memory):
 * A(w) = Kaa + i w beta Daa - w^2 Maa on the CSR pattern (indptr, indices) of the n active
 * unknowns; rhs(w) = -V (rK + i w beta rD - w^2 rM); Q = kts_a . x + V kts_z. */
typedef struct sawfem_operator {
  int n;
  const int64_t* indptr;
  const int* indices;
  const void* Kaa;
  const void* Maa;
  const void* Daa;        /* NULL without damping (then rD NULL too) */
  const void* rK;
  const void* rM;
  const void* rD;
  const void* kts_a;
  double kts_z_re, kts_z_im;
  double beta;
} sawfem_operator;

/* Bytes of the workspace sawfem_sweep_chunk needs for B frequencies in precision prec. */
int64_t sawfem_sweep_chunk_bytes(void* plan, int n, int B, int prec);

/* Q(w_b) of B frequencies (omega: host, rad/s): factor (prec 0 complex64 / 1 complex128),
 * right-hand sides, solve, and refinement against the FP64 operator until the estimated
 * relative error of each Q is <= ir_tol (at most max_ir corrections); with fallback, the
 * complex64 frequencies that do not converge are refactored and refined in complex128.  Q (host,
 * 2B doubles re/im), err (host, B: estimated relative error), flags (host, B: bit 0 refactored
 * in complex128, bit 1 not converged in the last precision tried).  Returns 0, 6 (done, some
 * frequencies refactored), 5 (a charge is not finite; outputs filled) or an error status.
 * workspace: device memory of at least sawfem_sweep_chunk_bytes(plan, n, B, prec) bytes.
 * Replaces FrequencySolver.solve of saw2d/sweep.py:335-364 for a whole chunk of frequencies. */
int sawfem_sweep_chunk(void* plan, void* stream, const sawfem_operator* op, int B, const double* omega, double V,
                       int prec, double ir_tol, int max_ir, int fallback, void* workspace, int64_t workspace_bytes,
                       double* Q, double* err, int* flags);

/* sawfem_sweep_chunk with factorization sharing (complex64): groups of `share` consecutive
 * frequencies are factored once, at the mean of their angular frequencies, each frequency is
 * refined against its own FP64 operator with those factors (the same stopping and stall rule),
 * and the solves take the vectors of a group together (sawfem_mf_solve_groups).  With fallback,
 * frequencies that do not converge are refined with their own complex64 factors (flags bit 2),
 * and those that still do not in complex128 (bit 0).  share = 1 is sawfem_sweep_chunk; prec 1
 * ignores share.  workspace: at least sawfem_sweep_chunk_share_bytes(plan, n, B, prec, share).
 * Replaces FrequencySolver.solve of saw2d/sweep.py:335-364 for a chunk (the reference factors
 * every frequency). */
int64_t sawfem_sweep_chunk_share_bytes(void* plan, int n, int B, int prec, int share);
int sawfem_sweep_chunk_share(void* plan, void* stream, const sawfem_operator* op, int B, const double* omega,
                             double V, int prec, int share, double ir_tol, int max_ir, int fallback, void* workspace,
                             int64_t workspace_bytes, double* Q, double* err, int* flags);

/* sawfem_adjoint_elem without the work argument: allocates nothing.  No counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The B * ne0 * NT complex128 element sums live
 * in a static device array of 64 MiB per GPU (held from library load until the process ends);
 * larger requests return 1.  Every call on one GPU uses that GPU's array: calls that overlap on
 * one GPU race. */
int sawfem_adjoint_elem_noalloc(void* stream, int ne0, const int* elems, const int* conn, const double* xy,
                                const int* elem_tag, int npml, const int* pml_tag, const int* pml_kind,
                                const double* pml_par, double omega_top, int ndu, int N, const int* pos_rep,
                                const unsigned char* zmask, double V, double scale, int B, int n, const void* x,
                                const void* lam, void* W);

#ifdef __cplusplus
}
#endif
#endif
