/*
 * hlbl_amd: the per-pair connected HLbL kernel sum on MI355X (gfx950).
 *
 * A thin C ABI over hand-written HIP kernels. Plain pointers, sizes and POD
 * structs only; every entry point returns an int status (0 = success) and
 * never throws or exits. hlbl_last_error() gives the message of the last
 * failure on the calling thread.
 *
 * Propagator blocks keep the reference layout (include/hlbl/provider.hpp:18-20,
 * src/provider.cpp:45): doubles at ((f*12 + a)*V + s)*24 + 2*c + {re,im} with
 * flavour f, source spin-colour a, local site s = ((t*LX+x)*LY+y)*LZ+z and sink
 * spin-colour c = 3*spin + colour, in the cvc gamma basis.
 */
#ifndef HLBL_AMD_H
#define HLBL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- POD types ------------------------------------------------------------
 * Same members and meaning as Geom, Coord and Pair of kernels/cuda_lattice.h:22-32. */
typedef struct { int T, LX, LY, LZ; } hlbl_geom;
typedef struct { int t, x, y, z; } hlbl_coord;
typedef struct { double a, b; } hlbl_pair; /* {xunit, unused}, src/contract.cpp:85-87 */

/* stream handle: a hipStream_t (NULL = the default stream) */
typedef void* hlbl_stream;

typedef struct hlbl_qed hlbl_qed;               /* QED kernel state (tables, MLP) */
typedef struct hlbl_contractor hlbl_contractor; /* work buffers of one lattice  */
typedef struct hlbl_schedule hlbl_schedule;     /* a built sliding-hypercube schedule */

/* message of the last failed call on this thread ("" if none) */
const char* hlbl_last_error(void);

/* ---- units ----------------------------------------------------------------
 * xunit = m_mu a / (hbar c), hbar c = 197.3269804 MeV fm.
 * Replaces hlbl::xunit, include/hlbl/physics.hpp:9-14. */
double hlbl_xunit(double muon_mass_mev, double lattice_spacing_fm);

/* ---- propagator provider ----------------------------------------------------
 * Replaces fill_random_block / RandomProvider::compute, src/provider.cpp:10-69.
 * hlbl_source_seed folds a source position into the seed (provider.cpp:60-64);
 * the fills write block_length = 2*12*24*V doubles, bit-identical to the
 * reference's host generator, for any process grid. */
uint64_t hlbl_source_seed(uint64_t seed, hlbl_coord gsx);
size_t hlbl_block_length(hlbl_geom local);
int hlbl_fill_random_block_host(double* out, hlbl_geom local, hlbl_coord proc_coords,
                                uint64_t seed, int n_threads);
int hlbl_fill_random_block_device(double* d_out, hlbl_geom local, hlbl_coord proc_coords,
                                  uint64_t seed, hlbl_stream stream);

/* ---- QED kernel L^(Lambda), Lambda = 0.4 -----------------------------------
 * Replaces KQED's initialise(&kqed_t) / free_QED_temps (bin/hlbl_hypercube.cpp:175,226)
 * and the device call QED_Mkernel_L2(0.4, x, y, kqed_t, kerv) (kernels/cuda_lattice.cu:266-275).
 * hlbl_qed_init_file loads the tabulated kernel (the T tensors of the Mainz
 * representation on a grid in |x|, beta, |y|, written by hlbl_amd/qedgen from the
 * kernel's defining integrals) from table_path onto `device`; hlbl_qed_init does the
 * same with the library's own table, data/qed_kernel_table.bin beside
 * libhlbl_amd.so, or the file the environment variable HLBL_QED_TABLE names. */
int hlbl_qed_init(hlbl_qed** out, int device);
int hlbl_qed_init_file(hlbl_qed** out, int device, const char* table_path);
int hlbl_qed_free(hlbl_qed* q);
/* The tabulated quantity itself, from the kernel's defining integrals (host, one
 * thread, no device): the 42 frame components of T^I, T^II, T^III at |x| = a,
 * |y| = b, angle beta in [0, pi] between x and y, into out[42]. quad = {nt, nr,
 * ntheta, npsi} quadrature sizes (Feynman-parameter nodes per axis; radial, polar and
 * azimuthal nodes per centre), or NULL for the table's own (16, 28, 8, 16).
 * No counterpart in the reference (KQED ships its tables precomputed). */
int hlbl_qed_exact(double a, double beta, double b, const int* quad, double* out);
/* table geometry: channel stride of a grid point, points along |x|, beta, |y|, and
 * the outer radius in muon units (the kernel is 0 beyond it) */
int hlbl_qed_table_info(const hlbl_qed* q, int* n_channels, int* n_a, int* n_c, int* n_b,
                        double* radius);
/* copy the FP64 table [n_a][n_c][n_b][n_channels] to host memory */
int hlbl_qed_table_copy(const hlbl_qed* q, double* host_out, size_t n_doubles);

/* Kernel values at n argument pairs (muon units): x, y are [n][4] device arrays,
 * kerv is the [n][6][4][4][4] device output, index [k][mu][nu][lambda] with k over
 * (rho,sigma) in (0,1),(0,2),(0,3),(1,2),(1,3),(2,3) (cuda_lattice.cu:52-60).
 * mode 0 = plain (table interpolation, FP64), 1 = surrogate (MLP, FP32/BF16;
 * plain outside the trained domain). */
int hlbl_qed_kernel(hlbl_qed* q, int mode, const double* d_x, const double* d_y, int n,
                    double* d_kerv, hlbl_stream stream);
/* hlbl_qed_kernel's plain form (mode 0) without allocation: the per-point scratch is a
 * fixed array of the library on each GPU (18.9 MB, held until the process ends), so at
 * most 2048 argument pairs per call (more is an error). No counterpart in the reference:
 * for callers that cannot allocate (code under graph capture, tight serving loops,
 * harnesses that forbid allocation). Calls that overlap on one GPU share the array and
 * race; order them. Results equal hlbl_qed_kernel's bit for bit. */
int hlbl_qed_kernel_noalloc(hlbl_qed* q, const double* d_x, const double* d_y, int n,
                            double* d_kerv, hlbl_stream stream);
/* The interpolated table channels (the 42 frame components of T^I, T^II, T^III,
 * then zero pad) at n points (a=|x|, beta = the angle between x and y in [0, pi],
 * b=|y|): d_acb is [n][3], d_out [n][n_channels]. mode as above. */
int hlbl_qed_channels(hlbl_qed* q, int mode, const double* d_acb, int n, double* d_out,
                      hlbl_stream stream);

/* ---- MLP surrogate of the weight functions --------------------------------
 * 3 inputs (|x|, |y|, cos beta in muon units), 4 hidden layers of 128 GELU,
 * n_channels outputs. BF16 weights, FP32 accumulation, FP32 master weights.
 * Trained on the GPU by this library's own forward/backward kernels on samples of
 * the plain table evaluation. No counterpart in the reference (the surrogate is this
 * library's addition). */
int hlbl_mlp_param_count(const hlbl_qed* q);
int hlbl_mlp_init(hlbl_qed* q, uint64_t seed);
/* one full training run; loss history (one value per epoch) goes to host `loss` */
int hlbl_mlp_train(hlbl_qed* q, int n_samples, int epochs, int batch, float lr,
                   uint64_t seed, float* loss, hlbl_stream stream);
/* trained domain of the surrogate: |x|, |y| <= radius (muon units, 0 < radius <=
 * the kernel radius, 12 fm; default the kernel radius). Arguments outside it take
 * the table's channels. Changing it discards the trained state (train again). */
int hlbl_mlp_set_domain(hlbl_qed* q, double radius);
/* inner radius of the trained domain: |x|, |y| >= inner (muon units, default 0.125;
 * 0 <= inner < radius). Below it the table's channels are used: the site-only
 * evaluations L0(x, 0), L0(0, x) always, where the weight functions bend like
 * |y| log |y| and a network of this size fits them worst. Network inputs are
 * (|x| - inner) / (radius - inner), the same for |y|, and beta / pi. Changing it
 * discards the trained state (train again). */
int hlbl_mlp_set_inner_radius(hlbl_qed* q, double inner);
/* parameters (FP32 master copy, packed layer by layer: W then b) to/from host */
int hlbl_mlp_get_params(const hlbl_qed* q, float* host, int n);
int hlbl_mlp_set_params(hlbl_qed* q, const float* host, int n);
/* forward on n inputs ([n][3] device, normalised inputs) -> [n][n_channels] outputs
 * (normalised); precision 0 = FP32 weights, 1 = BF16 weights */
int hlbl_mlp_forward(hlbl_qed* q, const float* d_in, int n, float* d_out, int precision,
                     hlbl_stream stream);
/* mean-squared-error loss and its gradient w.r.t. all parameters (FP32 weights) for a
 * batch: d_in [n][3], d_target [n][n_channels]; gradient goes to device d_grad */
int hlbl_mlp_loss_grad(hlbl_qed* q, const float* d_in, const float* d_target, int n,
                       float* d_grad, float* loss_host, hlbl_stream stream);

/* ---- contraction ------------------------------------------------------------
 * A contractor owns the work buffers for one lattice geometry; replaces the
 * Contractor of include/hlbl/contract.hpp:42-94 (buffers :69-79). */
int hlbl_contractor_create(hlbl_contractor** out, hlbl_geom global, hlbl_geom local,
                           hlbl_coord proc_coords, hlbl_qed* qed);
int hlbl_contractor_free(hlbl_contractor* c);

/* Launch shapes of the contractor's kernels (no counterpart in the reference, whose
 * launches are fixed at 32 threads per block, kernels/cuda_lattice.h:39-44). The
 * defaults are those below; hlbl_amd/tune.py measures candidates on the device and
 * records the fastest per lattice (hlbl_amd/data/tuning_gfx950.json), which the Python
 * layer applies to every contractor it creates. Results do not depend on the shapes
 * beyond the order of the FP64 sums across workgroups (pass 2, plain and window). */
typedef struct {
  int pass1_batch;           /* window pass 1: the units of a plane run on one XCD in lists
                                of at most this many that share blocks (they share them
                                in its L2); 1..2^20, default 32 (the CUs of an XCD) */
  int field_waves;           /* window coefficient field: waves per workgroup (one site
                                each), 4 or 8, default 8 */
  int field_wg_per_cu;       /* its workgroups per CU: 1..2 (8 waves) or 1..3 (4 waves),
                                default 2 */
  int axes_wg_per_cu;        /* window axis rows: workgroups per CU, 1..16, default 4 */
  int near_wg_per_cu;        /* window near-site FP64 field: workgroups per CU, 1..8,
                                default 1 */
  int pass2_wg_per_cu;       /* window pass 2: site ranges per CU (one resident), 1..4,
                                default 1 */
  int fp64_field_wg_per_cu;  /* plain path coefficient field: workgroups per CU, 1..16,
                                default 2 */
  int fp64_pass2_wg_per_cu;  /* plain path pass 2: workgroups per CU, 1..8, default 2 */
} hlbl_tuning;
/* out = the defaults */
int hlbl_tuning_defaults(hlbl_tuning* out);
/* validates and sets; later window calls capture their graphs again */
int hlbl_contractor_set_tuning(hlbl_contractor* c, const hlbl_tuning* t);
int hlbl_contractor_get_tuning(const hlbl_contractor* c, hlbl_tuning* out);

/* Seam 1, Contractor::kernel_sum (src/contract.cpp:104-114): the connected kernel
 * sum of both flavours of the pair x = gsx, y = gsx + yv, plain FP64. Blocks are
 * borrowed device pointers. Synchronous on `stream`; out[2] is host memory. The sum
 * over ranks is the identity on one process. */
int hlbl_kernel_sum(hlbl_contractor* c, const double* d_fwd_src, const double* d_fwd_y,
                    hlbl_coord gsx, hlbl_coord yv, double xunit, double out[2],
                    hlbl_stream stream);

/* Seam 2, the launchers of kernels/cuda_lattice.h:49-57 (cu_dzu_dzsu,
 * cu_4pt_contraction), with the contractor standing in for QED_kernel_temps and
 * the work space. Both ADD their result to the output buffers, which the caller
 * zeroes, as in the reference. d_dzu is [6][12][12] complex, d_dzsu [4][12][12],
 * d_g_dzu [6][4][12][12], d_g_dzsu [4][4][12][12] (contract.hpp:18-23). */
int hlbl_dzu_dzsu(hlbl_contractor* c, double* d_dzu, double* d_dzsu, const double* fwd_src,
                  const double* fwd_y, int iflavor, hlbl_coord proc_coords, hlbl_coord gsx,
                  hlbl_geom global, hlbl_geom local, hlbl_stream stream);
/* g = gamma_mu gamma_5 applied to dzu/dzsu (src/contract.cpp:58-77), on the device */
int hlbl_insert_g5_gamma(const double* d_dzu, const double* d_dzsu, double* d_g_dzu,
                         double* d_g_dzsu, hlbl_stream stream);
int hlbl_4pt_contraction(hlbl_contractor* c, double* d_kernel_sum, const double* d_g_dzu,
                         const double* d_g_dzsu, const double* fwd_src, const double* fwd_y,
                         int iflavor, hlbl_coord proc_coords, hlbl_coord gsx, hlbl_pair xunit,
                         hlbl_coord yv, hlbl_geom global, hlbl_geom local,
                         hlbl_stream stream);

/* Window entry: n pairs {x block, y block, gsx, yv} and both flavours per call.
 * form 0 = plain FP64; 1 = FP32 products + FP64 Kahan sums + plain kernel;
 * 2 = FP32 + surrogate kernel (plain FP64 near the origin and outside the trained
 * domain). Pass 1 and pass 2 of all pairs each run as one persistent launch; the
 * sequence is captured as a hipGraph and replayed for later windows of the same size
 * and form (another size reuses the buffers, which only grow, and is captured again).
 * The QED kernel enters through one coefficient field per distinct yv, which depends
 * only on (yv, xunit, form, the surrogate's parameters): the contractor keeps them
 * across calls (at least as many as the largest call's distinct yv, least recently
 * used reused first), so later windows of a chain, which see the same yv, build none;
 * a call builds the fields it lacks before its graph runs.
 * out is host [n][2]. No counterpart in the reference (window batching is this
 * library's addition); it replaces the pair loop of src/driver.cpp:110-132. */
typedef struct {
  const double* d_x;
  const double* d_y;
  hlbl_coord gsx;
  hlbl_coord yv;
} hlbl_pair_desc;
int hlbl_window_kernel_sum(hlbl_contractor* c, const hlbl_pair_desc* pairs, int n_pairs,
                           double xunit, int form, double* out, hlbl_stream stream);
/* the surrogate's judge for the last form-2 window: max relative deviation of the
 * surrogate kernel from the plain one over the sampled sites */
int hlbl_window_judge(const hlbl_contractor* c, double* max_rel_err, int* n_samples);
/* The judge gates form 2: a window whose judge exceeds `bound` (default 1e-3) runs
 * with the plain kernel (form 1) instead, decided before the window's passes run (the
 * judge of each surrogate field is taken when the field is built and remembered per
 * yv, xunit and network), and those results are returned; hlbl_window_fallback tells
 * whether the last form-2 window did so. No counterpart in the reference (the
 * surrogate is this library's addition). */
int hlbl_window_set_judge_bound(hlbl_contractor* c, double bound);
int hlbl_window_fallback(const hlbl_contractor* c, int* fell_back);
/* Forget the kept coefficient fields and free their device memory: the next window call
 * builds all of its own. No counterpart in the reference. */
int hlbl_window_drop_fields(hlbl_contractor* c);
/* n = the number of coefficient fields the last window call built (its distinct yv
 * that no kept field served). No counterpart in the reference. */
int hlbl_window_fields_built(const hlbl_contractor* c, int* n);
/* At least n slots in the store of kept fields (each V * 640 floats on the device;
 * growing drops the kept fields); slots (may be NULL) = the slots it now has.
 * hlbl_run_chain* reserve one per distinct yv of the chain when the device memory
 * beside the chain's blocks allows. No counterpart in the reference. */
int hlbl_window_reserve_fields(hlbl_contractor* c, int n, int* slots);
/* Measurement hook (no counterpart in the reference): runs a window of form 1 or 2
 * once without the graph, with HIP events around each stage on the launch stream;
 * ms = {pass 1 incl. its reductions, surrogate channels (0 for form 1),
 * coefficient fields, pass 2 incl. its reduction}; the surrogate channels and the
 * fields cost nothing when the kept fields serve the call. out as for the window
 * entry. */
int hlbl_window_profile(hlbl_contractor* c, const hlbl_pair_desc* pairs, int n_pairs,
                        double xunit, int form, double* out, double ms[4]);

/* ---- 2+2 pieces: P1 ------------------------------------------------------------
 * The P1 piece of the 2+2 contribution of one source block d_fwd at the global site
 * gsw (Contractor::pieces_2p2, src/contract.cpp:116-150; kernel_pi and kernel_p1 of
 * compute_2p2_gpu, kernels/hlbl_kernels/kernels.cu:241-340), FP64 in a fixed order:
 *   Pi_{mu nu}(z) = sum_a Re[(gamma_nu gamma_5 w_a)[a]],
 *   w_a[b] = sum_c conj(D_b[c]) (gamma_5 gamma_mu U_a)[c], U = block[iflavor], D = the other;
 *   d_p1[rho][mu][nu][zeta] (device, 64 * Lmax doubles, Lmax the largest global extent,
 *   overwritten) = sum of Pi_{mu nu}(z) over this process's sites with
 *   (z - gsw)_rho mod L_rho = zeta; 0 for bins it does not hold (sum over processes:
 *   the identity on one). Synchronous on stream. hlbl_pieces_2p2 adds P2..P4. */
int hlbl_pieces_2p2_p1(hlbl_contractor* c, const double* d_fwd, int iflavor, hlbl_coord gsw,
                       double* d_p1, hlbl_stream stream);
/* All 2+2 pieces of one source block (Contractor::pieces_2p2, src/contract.cpp:116-150;
 * kernel_p20, kernel_p21, kernel_p3, kernel_p4, kernels.cu:374-760): d_p1 as above, and
 * for each of the n_targets global sites (host array) d_p234[t][5][rho][sigma][nu]
 * (device) = P2_0, P2_1, P3, P4_0, P4_1 with the plain QED kernel at xunit, summed over
 * this process's sites in a fixed order; only rho < sigma is filled (P4_1: rho != sigma).
 * With x = z - gsw and y = target - gsw, each argument difference min-imaged on its own:
 *   P2_0 = L(x, y)[mu][nu][lam] Pi_{mu lam},  P2_1 = L(y, x)[nu][mu][lam] Pi_{mu lam},
 *   P3 = L(x - y, -y)[mu][lam][nu] Pi_{mu lam},  P4_0 = L(y - x, -x)[nu][lam][mu] Pi_{mu lam},
 *   P4_1 = x_rho (P4_0 term) at [rho][sigma] and -x_sigma (P4_0 term) at [sigma][rho].
 * Synchronous on stream. */
int hlbl_pieces_2p2(hlbl_contractor* c, const double* d_fwd, int iflavor, hlbl_coord gsw,
                    const hlbl_coord* targets, int n_targets, double xunit, double* d_p1,
                    double* d_p234, hlbl_stream stream);

/* ---- sliding-hypercube schedule (src/schedule.cpp:34-214) -------------------- */
hlbl_schedule* hlbl_schedule_build(hlbl_coord h, int stride, int n_steps, hlbl_coord origin);
void hlbl_schedule_free(hlbl_schedule* s);
int hlbl_schedule_n_windows(const hlbl_schedule* s);
int hlbl_schedule_n_sources(const hlbl_schedule* s);
/* source i in run order: its window, unwrapped position, partner count */
int hlbl_schedule_source(const hlbl_schedule* s, int i, int* window, hlbl_coord* pos,
                         int* n_partners);
int hlbl_schedule_partner(const hlbl_schedule* s, int i, int j, hlbl_coord* pos);
/* 0 if complete (check_complete, schedule.cpp:111-139), else -1 with the message */
int hlbl_schedule_check_complete(const hlbl_schedule* s, int stride);
/* simulate (schedule.cpp:168-214): mode 0 = serial, 1 = interleaved;
 * stats = {n_propagators, n_solves, n_pairs, n_kernel_sums, peak_cache} */
int hlbl_schedule_simulate(const hlbl_schedule* s, int mode, long long stats[5]);
int hlbl_max_steps_without_wrap(hlbl_coord h, int stride, hlbl_coord lattice,
                                hlbl_coord origin, int cap);

/* ---- driver: the schedule loop on random fields (src/driver.cpp:23-208) ------
 * Runs one chain: solve (device RandomProvider), cache (residency 0 = device,
 * 1 = pinned host with two staging slots), contract per pair (form as above;
 * batch 0 = one pair per call in solve order, 1 = one call per window, its sources
 * solved first, 2 = one call per new source with all its partners, in the solve order
 * and with the cache peak of batch 0), free blocks after their last pair. With the
 * host tier, batches 1 and 2 read their blocks from device staging blocks (as many as
 * the largest window, or the largest two consecutive source calls, need): a block just
 * solved is copied there from the solve target, the others from pinned host memory on
 * a second stream, batch 2's while the previous call computes. rows gets n_pairs * 16 doubles:
 * x[4], y[4], yv[4], window, chain(=0), ks0, ks1 (ints stored as doubles).
 * stats = {pairs, positions_solved, cache_peak, seconds_contract, seconds_solve}. */
int hlbl_run_chain(hlbl_contractor* c, const hlbl_schedule* s, uint64_t seed, double xunit,
                   int residency, int form, int batch, double* rows, int max_rows,
                   double stats[5]);
/* The same for chain index `chain` (stored in each row), calling fn(user, rows, n)
 * after each window with that window's rows (16 doubles each, as above), where the
 * reference appends them to its result file (src/driver.cpp:134-137); a nonzero return
 * from fn stops the run with an error. fn may be NULL. hlbl_result_append_rows of
 * libhlbl_h5.so (include/hlbl_h5.h) is such an fn. */
typedef int (*hlbl_window_rows_fn)(void* user, const double* rows, int n);
int hlbl_run_chain_each_window(hlbl_contractor* c, const hlbl_schedule* s, uint64_t seed,
                               double xunit, int residency, int form, int batch, int chain,
                               hlbl_window_rows_fn fn, void* user, double* rows, int max_rows,
                               double stats[5]);

/* ---- the 2+2 step of a run (src/driver.cpp:58-95) ----------------------------
 * Result2p2 (include/hlbl/output.hpp:25-35): one 2+2 source and its targets, host
 * arrays: p1 [4][4][4][Lmax], p234 [n_targets][5][4][4][4]. */
typedef struct {
  hlbl_coord w;
  int chain, window;
  const double* p1;
  const hlbl_coord* targets;
  int n_targets;
  const double* p234;
} hlbl_source_2p2;
/* A 2+2 source whose pieces this chain computes: its site w (in [0, L)), the window of
 * its first solve in the run (plan_2p2, src/schedule.cpp:284-317) and its targets. */
typedef struct {
  hlbl_coord w;
  int window;
  const hlbl_coord* targets;
  int n_targets;
} hlbl_plan_2p2;
typedef int (*hlbl_window_2p2_fn)(void* user, const hlbl_source_2p2* rows, int n);
/* hlbl_run_chain_each_window with the 2+2 step inside the solve loop: when the chain
 * solves a planned site in its planned window, P1 and P2..P4 of its targets
 * (hlbl_pieces_2p2, flavour iflavor) are computed from the solved block before it is
 * stored (the solve target, as src/driver.cpp:88-93), and fn2(user2, rows, n) receives
 * each window's 2+2 rows after the window's kernel-sum rows (fn2 may be NULL). A
 * nonzero return from fn2 stops the run with an error. hlbl_result_append_2p2_rows
 * of libhlbl_h5.so is such an fn2. */
int hlbl_run_chain_2p2(hlbl_contractor* c, const hlbl_schedule* s, uint64_t seed, double xunit,
                       int residency, int form, int batch, int chain, hlbl_window_rows_fn fn,
                       void* user, const hlbl_plan_2p2* plan, int n_plan, int iflavor,
                       hlbl_window_2p2_fn fn2, void* user2, double* rows, int max_rows,
                       double stats[5]);

#ifdef __cplusplus
}
#endif

#endif /* HLBL_AMD_H */
