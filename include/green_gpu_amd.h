/*
 * C ABI of the gfx950 HF-exchange / GW self-energy library.
 *
 * This is the boundary green-mbpt's custom-kernel plugin API sits on.  The
 * reference exposes the same path as C++ objects behind three factory
 * functions (src/green/gpu/gpu_factory.h:47-62 custom_hf_kernel, :78-89
 * custom_gw_kernel, :95-102 custom_kernel_parameters) whose kernels own all
 * device memory for the duration of one solve (hf_gpu_kernel.cpp:66-104,
 * gw_gpu_kernel.cpp:163-205).  Here the same work is split into plans that
 * own their device memory across calls; include/green/gpu/gpu_factory.h
 * rebuilds the reference's C++ signatures on top of these functions.
 *
 * Conventions
 *   - complex numbers are interleaved (re, im) doubles, i.e. std::complex<double>;
 *   - every tensor is row-major with the reference's host layout;
 *   - functions return 0 on success or a ggpu_status code; the message of the
 *     last failure on the calling thread is returned by ggpu_last_error();
 *   - plans are not thread-safe: one thread per plan.
 */
#ifndef GREEN_GPU_AMD_H
#define GREEN_GPU_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  GGPU_OK = 0,
  GGPU_ERR_ARGUMENT = 1,             /* reference: std::runtime_error / std::logic_error on bad input */
  GGPU_ERR_DEVICE = 2,               /* HIP runtime failure                                            */
  GGPU_ERR_MEMORY = 3,               /* device or pinned allocation failed (cugw_qkpt.cu:39-75)       */
  GGPU_ERR_NOT_POSITIVE_DEFINITE = 4,/* 1 - P0(Omega) not HPD; reference aborts (cu_compute_Pq.cu:24-43) */
  GGPU_ERR_STATE = 5,                /* call out of order (e.g. iterate before integrals are loaded)  */
  GGPU_ERR_SINGULAR = 6              /* LU solver: exactly zero pivot in 1 - P0(Omega) (getrf info > 0;
                                        reference aborts, cu_compute_Pq.cu:24-43)                    */
} ggpu_status;

/* Linear solver for the Dyson equation, same order as LinearSolverType
 * (common_defs.h:137-139). */
typedef enum { GGPU_SOLVER_CHOLESKY = 0, GGPU_SOLVER_LU = 1 } ggpu_linear_solver;

/* Arithmetic of the G.V, P0 and Sigma contractions and the IR transforms.
 * The Dyson factorisation and the Hermitisation of P0 always run in FP64. */
typedef enum {
  GGPU_PRECISION_FP64 = 0,    /* reference "double" path (gw_gpu_kernel.cpp:278-284, _sp = false) */
  GGPU_PRECISION_FP32 = 1,    /* complex64 contractions (reference _sp = P_sp && Sigma_sp)        */
  GGPU_PRECISION_BF16X3 = 2,  /* complex64 data, split-bf16 x3 products, FP32 accumulation        */
  GGPU_PRECISION_INT8X5 = 3,  /* complex128 data; FP64-accurate products, each call site either
                                 emulated on the int8 matrix cores (five signed 7-bit slices per
                                 element, one power-of-two scale per row / column, exact int32
                                 slice products up to slice diagonal 5, FP64 recombination; ~1e-10
                                 relative per product) or run on the FP64 kernel, whichever the
                                 autotuner (or ggpu_*_set_variants) picks; built-in: emulation at
                                 the Sigma P.Y1 product and the P0 contraction.  No counterpart in
                                 the reference                                                   */
  GGPU_PRECISION_INT8X4 = 4   /* as INT8X5 with four slices per element (slice diagonals up to 4,
                                 13 slice products instead of 19): ~2^-27 of the operand scale
                                 per product, fewer int8 products and slice bytes; the autotuner
                                 offers four slices at the P0 contraction and the Sigma P.Y1
                                 product only, five slices (variant 201) and the FP64 kernel at
                                 every site.  No counterpart in the reference                    */
} ggpu_precision;

/* Where a plan keeps the stored integrals.  HOST: in pinned host memory,
 * streamed pair batch by pair batch into a double-buffered device staging
 * area on a copy stream of the plan's own, each batch's copy overlapping the
 * previous batch's compute (the reference's low-GPU-memory mode uploads every
 * (k, q) pair, cugw_qkpt.cu:115-160).  DEVICE (default): loaded once into HBM.
 * GENERATED: nothing is stored; each batch of synthetic pairs (the values
 * *_synth_integrals defines) is generated on the device into the staging area
 * when it is needed, so systems whose integrals fit neither HBM nor host
 * memory can be run; only *_synth_integrals fills such a plan.
 * READER: nothing is stored either; each batch's stored pairs are fetched
 * through the caller's ggpu_pair_reader (*_set_integral_reader) into pinned
 * host staging and copied like HOST batches, so real integrals larger than
 * host memory stream from disk: the reference's low-CPU-memory "chunks"
 * mode, df_integral_t::read_integrals (df_integral_t.h:61-84) reading the
 * chunk of each (k1, k2) pair on demand (gpu_kernel.h:47-51). */
typedef enum {
  GGPU_INTEGRALS_DEVICE = 0,
  GGPU_INTEGRALS_HOST = 1,
  GGPU_INTEGRALS_GENERATED = 2,
  GGPU_INTEGRALS_READER = 3
} ggpu_integral_residence;

/* Fills one stored k-pair (stored position pos, 0 <= pos < num_kpair_stored)
 * into dst: NQ * nao * nao complex128 as interleaved doubles, the layout of
 * one entry of *_load_integrals' array.  Returns 0 on success; any other
 * value fails the calling iterate / exchange with GGPU_ERR_ARGUMENT.  Called
 * on the thread that runs the plan, one pair batch ahead of the device. */
typedef int (*ggpu_pair_reader)(void* user, int64_t pos, double* dst);

/*
 * Symmetry and integral-storage tables: the plain data green-symmetry
 * provides (cu_symmetry.h:41-66, built at gw_gpu_kernel.cpp:34-91) plus the
 * k-pair tables the integral reader uses (df_integral_t.h:68-86).
 * Integer arrays are int64.  Optional members may be NULL:
 *   U_k == NULL      -> identity AO representation for every k
 *   U_q == NULL      -> identity auxiliary representation for every q
 *   conj_kpair/trans_kpair/irre_pos == NULL -> every pair k1 >= k2 stored at
 *                       its pair index k1(k1+1)/2 + k2
 */
typedef struct {
  int32_t nk, ink, nq, inq;
  const int64_t* k_full_to_reduced;  /* [nk]                                   */
  const int64_t* k_reduced_to_full;  /* [ink]                                  */
  const int64_t* k_tr_conj;          /* [nk]  non-zero: time-reversal image    */
  const int64_t* k_star_offsets;     /* [ink+1]                                */
  const int64_t* k_star_members;     /* [k_star_offsets[ink]] full-BZ k        */
  const int64_t* q_full_to_reduced;  /* [nq]                                   */
  const int64_t* q_reduced_to_full;  /* [inq]                                  */
  const int64_t* q_tr_conj;          /* [nq]                                   */
  const int64_t* q_star_offsets;     /* [inq+1]                                */
  const int64_t* q_star_members;     /* [q_star_offsets[inq]]                  */
  const int64_t* k1_from_k2q;        /* [nk*nq]  k1 = k2 + q, index k2*nq + q  */
  const int64_t* k2_from_k1q;        /* [nk*nq]  k2 = k1 - q, index k1*nq + q  */
  const double* U_k;                 /* [nk][nso][nso] complex, or NULL        */
  const double* U_q;                 /* [nq][NQ][NQ] complex, or NULL          */
  const int64_t* conj_kpair;         /* [nk(nk+1)/2] or NULL                   */
  const int64_t* trans_kpair;        /* [nk(nk+1)/2] or NULL                   */
  const int64_t* irre_pos;           /* [nk(nk+1)/2] stored position, -1 = not stored, or NULL */
  int64_t num_kpair_stored;          /* number of stored pairs                 */
} ggpu_symmetry;

typedef struct ggpu_hf_plan ggpu_hf_plan;
typedef struct ggpu_gw_plan ggpu_gw_plan;
typedef struct ggpu_gw_batch ggpu_gw_batch;

/* Message of the last failure on this thread ("" if none). */
const char* ggpu_last_error(void);

/* Number of visible GPUs (reference: check_for_cuda, cuda_check.cpp:29-47). */
int ggpu_device_count(int* count);

/* Bytes of device memory currently free / total on the current device. */
int ggpu_device_memory(size_t* free_bytes, size_t* total_bytes);

/* ------------------------------------------------------------------------
 * HF exchange  (replaces scalar_hf_gpu_kernel::compute_exchange_selfenergy,
 * hf_gpu_kernel.cpp:106-139, x2c_hf_gpu_kernel, hf_gpu_kernel.cpp:229-338,
 * and cuhf_utils, cuhf_utils.cu:32-159)
 *
 * nso = nao: scalar plan, ns = 1 or 2.
 * nso = 2 nao: two-component (X2C) plan, ns = 1; density and Fock matrices
 * are nso x nso spinor matrices (custom_hf_kernel(X2C = nso != nao, ...),
 * gpu_factory.h:47-62).  The k-point tables must not contain time-reversal
 * images for X2C (their spinor rotation is not part of this library).
 * ------------------------------------------------------------------------ */
int ggpu_hf_plan_create(ggpu_hf_plan** plan, int nao, int nso, int ns, int NQ, const ggpu_symmetry* sym);

/* Choose where the integrals live (ggpu_integral_residence); call before
 * loading or synthesising them.  Results do not depend on the choice. */
int ggpu_hf_set_integral_residence(ggpu_hf_plan* plan, int residence);
/* Arithmetic of the two exchange products X = V D and K += X V^dagger
 * (ggpu_precision; default FP64).  FP32 and split-bf16 x3 run them on
 * complex64 copies of V and D with K's partial sums, the Hartree term J and
 * the Ewald term kept in FP64.  The reference's exchange is FP64 or complex64
 * as a whole (cuhf_utils<prec>, cuhf_utils.cu:32-159). */
int ggpu_hf_set_precision(ggpu_hf_plan* plan, int precision);
/* Launch variant of the exchange products (11 entries indexed by call site
 * as for ggpu_gw_site_stats; only site 9, HF, may be set, -1 elsewhere): in
 * the GGPU_PRECISION_INT8X5 / INT8X4 modes 200 runs them on the int8
 * emulation with the mode's slice count, 201 on the emulation with five
 * slices, and 83 (the built-in choice) on the FP64 kernel.  No counterpart
 * in the reference. */
int ggpu_hf_set_variants(ggpu_hf_plan* plan, const int* variants);
/* This device's share of the IBZ k-points: exchange and solve then compute
 * the rows ik = rank, rank + size, ... < ink only and leave the other rows
 * zero, so the results of shares 0 .. size-1 sum to the full K or Fock.  The
 * reference's per-device k loop, cuhf_utils.cu:140 (k_reduced_id =
 * devices_rank; k_reduced_id += devices_size), whose Fock rows are summed by
 * the allreduce of hf_gpu_kernel.cpp:93.  Default: rank 0 of 1. */
int ggpu_hf_set_k_share(ggpu_hf_plan* plan, int rank, int size);

/* Copy the stored integral pairs (num_kpair_stored x NQ x nao x nao complex,
 * the buffer df_integral_t::read_entire fills, df_integral_t.h:96-129) to the
 * device.  The host array is not referenced after the call. */
int ggpu_hf_load_integrals(ggpu_hf_plan* plan, const double* V_stored);
/* READER residence: the plan's integrals come from `reader` (with `user`
 * passed back) per pair batch; replaces *_load_integrals.  Sets the residence
 * to GGPU_INTEGRALS_READER. */
int ggpu_hf_set_integral_reader(ggpu_hf_plan* plan, ggpu_pair_reader reader, void* user);

/* Fill the device integral store with the synthetic counter-based generator
 * (Re, Im ~ N(0, sigma^2); diagonal pairs Hermitised).  Benchmarks only. */
int ggpu_hf_synth_integrals(ggpu_hf_plan* plan, uint64_t seed, double sigma);

/* K_s(k) for the IBZ k:  K = (px/nk) sum_{k2,Q} V^Q(k,k2) D_s(k2) V^Q(k,k2)^dagger,
 * px = -0.5 (ns = 1) or -1 (ns = 2).  dm: (ns, ink, nso, nso) on the IBZ (the
 * library unfolds it to the full BZ, hf_gpu_kernel.cpp:203-209).
 * K_out: (ns, ink, nso, nso), overwritten.
 * X2C: the aa, bb and ab blocks are the exchange of the matching density
 * blocks with px = -1 and K_ba = K_ab^dagger (cuda_common.cu:176-205). */
int ggpu_hf_exchange(ggpu_hf_plan* plan, const double* dm, double* K_out);

/* The exchange K of ggpu_hf_exchange on caller device buffers, allocating
 * nothing: dm_device and K_device are (ns, ink, nso, nso) complex in device
 * memory (K_device overwritten); the K partial sums of the split (k2, Q)
 * reduction live in a static per-GPU scratch array of 64 MiB (4 Mi complex
 * doubles) that the library holds until the process ends, and every other
 * buffer the call touches was allocated when the plan was created.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (graph capture, tight serving loops, allocation-free harnesses).
 * Limits: one partial slice (nps * nao^2 complex elements, nps = ns, or 3
 * for X2C) must fit the scratch, else GGPU_ERR_ARGUMENT; the integrals must
 * be resident in HBM (GGPU_INTEGRALS_DEVICE), else GGPU_ERR_STATE.  Calls
 * on one GPU share the scratch: calls that overlap on one GPU (several host
 * threads, or plans on different streams) race and must be serialised by
 * the caller.  Returns after the result is complete. */
int ggpu_hf_exchange_device_noalloc(ggpu_hf_plan* plan, const void* dm_device, void* K_device);

/* Full HF self-energy, as scalar_hf_gpu_kernel::solve returns it
 * (hf_gpu_kernel.cpp:66-104, without the MPI reduction):
 *   F = J + K - c * madelung * S D S,   c = 1 (ns = 2) or 0.5 (ns = 1),
 *   J(k) = sum_Q V^Q(k,k) (1/nk) sum_{s,k'} Tr[V^Q(k',k') D_s(k')]  (hf_gpu_kernel.cpp:141-186).
 * dm, S_k, F_out: (ns, ink, nso, nso) on the IBZ.
 * X2C (hf_gpu_kernel.cpp:253-327): J from D_aa + D_bb on both diagonal
 * blocks; Ewald madelung * S_aa D_xy S_aa on the aa, bb, ab blocks and its
 * adjoint on ba. */
int ggpu_hf_solve(ggpu_hf_plan* plan, const double* dm, const double* S_k, double madelung, double* F_out);

/* Two-electron HF energy of the last ggpu_hf_solve on this plan:
 *   E = (1/nk) sum_{s, k in IBZ} w_k (1/2) Re Tr[D_s(k) F_s(k)],  w_k = |star(k)|,
 * i.e. the full-Brillouin-zone sum (Tr[D F] is invariant under the k rotation
 * and only conjugated under time reversal).  F is the returned J + K - Ewald;
 * with ns = 1 the density is the spin-summed one, so the formula is the same;
 * for X2C the traces run over the nso x nso spinor matrices.  The traces
 * accumulate in FP64 on the device in every precision mode; with a k share
 * only the share's rows contribute (shares sum to the total).  The reference
 * returns only the Fock matrix (hf_gpu_kernel.cpp:66-104); the energy is
 * green-mbpt's.  GGPU_ERR_STATE before the first solve. */
int ggpu_hf_energy(ggpu_hf_plan* plan, double* energy);

void ggpu_hf_plan_destroy(ggpu_hf_plan* plan);

/* ------------------------------------------------------------------------
 * GW self-energy  (replaces scalar_gw_gpu_kernel::compute_gw_selfenergy,
 * gw_gpu_kernel.cpp:295-360, and cugw_utils::accumulate_gw_selfenergy_on_device,
 * cugw_utils.cu:103-202)
 * ------------------------------------------------------------------------ */
/* T_wt: (nw_b, nts-2) complex  (transformer_t::Tnt_BF, cugw_qpt.cu:329-339)
 * T_tw: (nts, nw_b)   complex  (transformer_t::Ttn_FB, cugw_qpt.cu:341-358)
 * nts must be even and nts >= nw_b (gw_gpu_kernel.h:67, cugw_qpt.cu:79).
 * work_bytes: device workspace budget for pair batches (0 = automatic).
 * nso = nao: scalar plan (ns = 1 or 2).  nso = 2 nao: two-component plan
 * (ns = 1), whose contractions run on the pseudo-spins aa, bb, ab, ba with
 * G(k, beta - tau) spin-swapped (x2c_gw_gpu_kernel, gw_gpu_kernel.cpp:458-551,
 * cugw_qkpt.cu:381-403); no time-reversal k images. */
int ggpu_gw_plan_create(ggpu_gw_plan** plan, int nao, int nso, int ns, int NQ, int nts, int nw_b,
                        const ggpu_symmetry* sym, const double* T_wt, const double* T_tw,
                        int precision, int solver, size_t work_bytes);

int ggpu_gw_set_integral_residence(ggpu_gw_plan* plan, int residence);
/* Bytes of integrals copied host -> device by the last iterate (0 when resident in HBM). */
int ggpu_gw_streamed_bytes(ggpu_gw_plan* plan, double* bytes);
/* Bytes of integrals generated on the device by the last iterate (GENERATED residence). */
int ggpu_gw_generated_bytes(ggpu_gw_plan* plan, double* bytes);
/* Device time (ms) the plan's copy stream spent staging those integrals in
 * the last iterate: the host -> device copies (HOST residence) or the
 * generation kernels (GENERATED), an event pair around every pair batch;
 * bytes / time is the rate the staging runs at while it runs.  0 when the
 * integrals are resident in HBM. */
int ggpu_gw_staging_ms(ggpu_gw_plan* plan, double* ms);
/* This device's share of the IBZ q-points: iterations then run
 * iq = rank, rank + size, ... < inq only, and Sigma holds that share's
 * contribution (still divided by nk), so the Sigma of shares 0 .. size-1
 * sum to the full one.  The reference's per-device q loop,
 * cugw_utils.cu:115 (q_ibz_idx = devices_rank; q_ibz_idx += devices_size).
 * Default: rank 0 of 1. */
int ggpu_gw_set_q_share(ggpu_gw_plan* plan, int rank, int size);
int ggpu_gw_load_integrals(ggpu_gw_plan* plan, const double* V_stored);
int ggpu_gw_synth_integrals(ggpu_gw_plan* plan, uint64_t seed, double sigma);
/* READER residence for a GW plan (as ggpu_hf_set_integral_reader). */
int ggpu_gw_set_integral_reader(ggpu_gw_plan* plan, ggpu_pair_reader reader, void* user);

/* One GW iteration: Sigma(tau) from G(tau).
 * G, Sigma_out: (nts, ns, ink, nso, nso) complex, Sigma_out overwritten
 * (already divided by nk, as gw_gpu_kernel.cpp:188 leaves it). */
int ggpu_gw_iterate(ggpu_gw_plan* plan, const double* G, double* Sigma_out);

/* Same, with G and Sigma already in device memory (HBM-resident inputs). */
int ggpu_gw_iterate_device(ggpu_gw_plan* plan, const void* G_device, void* Sigma_device);

/* Several independent systems (plans of any sizes) in one call: each plan
 * runs its iteration on its own stream from its own host thread, so small
 * systems overlap on the device.  G_device[i] / Sigma_device[i] as for
 * ggpu_gw_iterate_device; a plan may appear only once.  On failure the
 * message names the first failing system.  No counterpart in the reference,
 * which runs one system per process (gw_gpu_kernel::solve,
 * gw_gpu_kernel.cpp:163-205, called once per iteration). */
int ggpu_gw_iterate_many(ggpu_gw_plan* const* plans, int n, const void* const* G_device, void* const* Sigma_device);

/* Several independent systems as ONE HIP graph launch: the batch captures
 * every plan's whole iteration (as ggpu_gw_set_graph does for one plan), each
 * on its plan's stream, forked from and joined to the batch's own stream, so
 * ggpu_gw_batch_run replays all systems concurrently with one launch.  G /
 * Sigma device buffers as for ggpu_gw_iterate_device, fixed at creation.
 * Every plan needs its integrals resident in HBM (GGPU_ERR_STATE otherwise);
 * a later change to a plan (q share, variants, autotune, integrals) makes the
 * next run capture again.  Plans must outlive the batch and must not be used
 * elsewhere while it runs.  ms (may be NULL): device time of the replay.
 * On a failed Dyson factorisation the message names the first failing
 * system.  No counterpart in the reference (one system per process). */
int ggpu_gw_batch_create(ggpu_gw_batch** batch, ggpu_gw_plan* const* plans, int n, const void* const* G_device,
                         void* const* Sigma_device);
int ggpu_gw_batch_run(ggpu_gw_batch* batch, double* ms);
/* Number of graph captures the batch has made. */
int ggpu_gw_batch_captures(ggpu_gw_batch* batch, int* captures);
void ggpu_gw_batch_destroy(ggpu_gw_batch* batch);

/* Intermediates of one IBZ q for error reports and tests:
 * which = 0: P0(tau)   (nts, NQ, NQ)   bare bubble, scaled by 1/nk and mirrored
 * which = 1: P(Omega)  (nw_b, NQ, NQ)  after the Dyson solve
 * which = 2: P(tau)    (nts, NQ, NQ)   dressed polarization */
int ggpu_gw_polarization(ggpu_gw_plan* plan, const double* G, int iq, int which, double* out);

/* The same intermediates, read back without recomputation, for the last q
 * the last ggpu_gw_iterate* (or ggpu_gw_polarization) processed: *iq is that
 * IBZ q.  With a pair-batch limit they are partial.  GGPU_ERR_STATE before
 * any q has run.  No counterpart in the reference (its buffers are private). */
int ggpu_gw_last_polarization(ggpu_gw_plan* plan, int which, int* iq, double* out);

/* Restrict every q of the next iterates to its first `batches` pair batches
 * of the P0 chain and of the Sigma chain (per star member); 0 (default)
 * runs all pairs.  A partial iteration for warm-up and profiling runs: Sigma
 * is then not the full self-energy.  No counterpart in the reference. */
int ggpu_gw_set_pair_batch_limit(ggpu_gw_plan* plan, int batches);

/* Wall time (ms, device events) of the phases of the last iterate:
 * [0] total, [1] G staging, [2] P0, [3] IR + Dyson, [4] Sigma, [5] Sigma staging. */
int ggpu_gw_last_timings(ggpu_gw_plan* plan, double* ms6);

/* Per-call-site timing of the GEMM launches of ggpu_gw_iterate*: a pair of
 * device events on the plan's stream around every launch (the reference
 * reports only an iteration-level FLOP estimate, gw_gpu_kernel.cpp:207-247).
 * Sites: 0 other, 1 P0 G.V, 2 P0 G.V^dagger, 3 P0 contraction, 4 Sigma V.G,
 * 5 Sigma P.Y1, 6 Sigma contraction, 7 IR transforms, 8 Dyson updates,
 * 9 HF, 10 symmetry rotations. */
int ggpu_gw_set_profiling(ggpu_gw_plan* plan, int enable);
/* Totals of the last profiled iteration for one site: device ms, algorithmic
 * FLOPs (8 per complex multiply-add), launch count. */
int ggpu_gw_site_stats(ggpu_gw_plan* plan, int site, double* ms, double* flops, int64_t* launches);

/* Graph replay (enable != 0): the next iterate records the iteration's
 * launches (G staging, every q's P0, IR, Dyson and Sigma chains, the Sigma
 * layout change) into one HIP graph and replays it; later iterates replay it
 * with one launch.  A new capture is made when G or Sigma change address
 * (ggpu_gw_iterate always uses the plan's own device copies) or after
 * ggpu_gw_set_q_share, *_set_variants, *_autotune, *_load_integrals,
 * *_synth_integrals or *_set_integral_residence.  With staged integrals
 * (HOST, GENERATED) the plan's copy stream joins the capture, so the staging
 * copies or generation kernels of every pair batch are nodes of the graph and
 * are replayed with it; the READER residence (host-side reads) and per-site
 * profiling are refused (GGPU_ERR_STATE).  The results are those of the
 * launches it replays, and ggpu_gw_last_timings reports only the total.  No counterpart in the
 * reference, which enqueues every launch from the host each iteration
 * (cugw_utils.cu:103-202). */
int ggpu_gw_set_graph(ggpu_gw_plan* plan, int enable);
/* Number of graph captures the plan has made. */
int ggpu_gw_graph_captures(ggpu_gw_plan* plan, int* captures);

/* Autotuner: pick the launch variant of each contraction call site of this
 * plan by timing every candidate on the plan's own shapes and data, inside the
 * iteration's own launch sequence (operands cold in the caches as they are in
 * an iteration): the P0 and Sigma chains of the first q of this plan's share,
 * `batches` pair batches per chain (and per star member), best of `rounds`
 * (1 .. 8).  Candidates (variant ids): the LDS k-depth and waves-per-SIMD
 * register allocation for FP64 / FP32 (8, 16, 83, 84); the workgroup tile for
 * split-bf16 (0, 1, 4 .. 9; 10 .. 12: 128 x 128 from operands split once into
 * bf16 planes in the plan's workspace, four, eight or sixteen waves; 13:
 * 128 x 192, eight waves); in the
 * int8 modes the emulation with the mode's slices (200), with five slices
 * (201, four-slice plans) or the FP64 kernel (83).  Within FP64, FP32 and
 * split-bf16 the choice changes no result (every variant sums in the same
 * order); in the int8 modes it changes the rounding within the mode's
 * bound.  The choice is kept by the
 * plan and used by every later call.  Nothing the caller sees is written:
 * the plan's own buffers hold the trial results, and the next iterate starts
 * from zero.  G / G_device as for ggpu_gw_iterate / ggpu_gw_iterate_device.
 * No counterpart in the reference, which leaves kernel choice to cuBLAS's
 * internal heuristics (cublas_routines_prec.cu:53-228). */
int ggpu_gw_autotune(ggpu_gw_plan* plan, const double* G, int rounds, int batches);
int ggpu_gw_autotune_device(ggpu_gw_plan* plan, const void* G_device, int rounds, int batches);
/* What the last autotune measured: *n candidates (their variant ids in
 * variants[0 .. min(cap, n) - 1]); ms[c * 11 + site]: best time of candidate
 * c at that call site (site numbers as for ggpu_gw_site_stats; negative:
 * not run); *wall_ms: wall time of the tuning.  Output arrays may be NULL. */
int ggpu_gw_autotune_report(ggpu_gw_plan* plan, int cap, int* n, int* variants, double* ms, double* wall_ms);
/* The plan's launch variant per call site (11 entries, -1: built-in choice),
 * to save a tuning and give it to another plan of the same shape and
 * precision.  set rejects a variant the plan's precision does not have or a
 * site that is not one of the six contraction sites (1 .. 6). */
int ggpu_gw_get_variants(ggpu_gw_plan* plan, int* variants);
int ggpu_gw_set_variants(ggpu_gw_plan* plan, const int* variants);

void ggpu_gw_plan_destroy(ggpu_gw_plan* plan);

#ifdef __cplusplus
}
#endif

#endif /* GREEN_GPU_AMD_H */
