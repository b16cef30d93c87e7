/* ghx — C ABI of the MI355X-native gauge+Higgs HMC molecular-dynamics path.
 *
 * One context holds one or more lattices ("replicas") of the same gauge group
 * SU(N), dimension D, representation and precision, resident in HBM on the
 * current HIP device. All calls return 0 (GHX_OK) or a negative GHX_E_* code;
 * ghx_last_error() gives the message. No exceptions cross this boundary.
 * A context is not re-entrant; separate contexts are independent.
 *
 * Field buffers exchanged with the host use the reference's AoS layouts:
 *   U   [(s*D+mu)*N*N + i*N + j] complex (re,im interleaved)  fields.hpp:13-16
 *   P   [(s*D+mu)*(N*N-1) + a]    real, generator order a       fields.hpp:47-50
 *   phi [s*d + k]                 complex                       scalar_field.hpp:12-18
 * with s the site index of that replica (direction 0 fastest, geometry.hpp:12-53).
 */
#ifndef GHX_H
#define GHX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ghx_ctx ghx_ctx;

enum {
  GHX_OK = 0,
  GHX_E_INVALID = -1,     /* bad argument or state */
  GHX_E_HIP = -2,         /* HIP runtime failure */
  GHX_E_NAN = -3,         /* non-finite Hamiltonian */
  GHX_E_UNSUPPORTED = -4, /* (N, D, representation) combination not compiled */
  GHX_E_REP = -5          /* invalid representation spec (GeneralRep ctor errors) */
};
enum { GHX_FP64 = 0, GHX_FP32MD = 1 };        /* FP32MD: U,P,phi,pi,forces in FP32; H, dH in FP64 */
enum { GHX_LEAPFROG = 0, GHX_OMELYAN = 1 };   /* hmc.hpp:10 Integrator */
enum { GHX_COLD = 0, GHX_HOT = 1 };           /* start modes, fields.hpp:21-34 */

/* Create a context on the current HIP device. L holds n_replicas*D extents.
 * Replaces the construction of GaugeHMC / GaugeHiggsHMC (hmc.hpp:29-30,
 * gauge_higgs_hmc.hpp:45-50) and of ReplicaTempering's replica set
 * (replica_tempering.hpp:97-118). */
int ghx_create(ghx_ctx** out, int D, int N, int n_replicas, const int* L, int precision);
int ghx_destroy(ghx_ctx* ctx);
const char* ghx_last_error(const ghx_ctx* ctx);

/* Representation of the scalar: "" or "none" (pure gauge, GaugeHMC semantics),
 * "fund", "adj", or Young rows "r1,r2,..." with optional ":real"
 * (hmc_higgs.cpp:23-34 make_rep; rep_general.hpp:54-84 GeneralRep ctor). */
int ghx_set_rep(ghx_ctx* ctx, const char* spec, int* d_out);
/* d, N-ality, real flag, tensor dimension N^n (0 for fund/adj/none)
 * (representation.hpp:14-17). */
int ghx_rep_info(const ghx_ctx* ctx, int* d, int* nality, int* real_rep, long* dimT);
/* T^a_R, (N*N-1)*d*d complex (representation.hpp:17). */
int ghx_rep_generators(const ghx_ctx* ctx, double* T);
/* Orthonormal real basis W, d*dimT (rep_general.hpp:40). */
int ghx_rep_basis(const ghx_ctx* ctx, double* W);

/* Couplings of one replica (GaugeHiggsHMC public fields, gauge_higgs_hmc.hpp:24-30). */
int ghx_set_couplings(ghx_ctx* ctx, int replica, double beta, double kappa, double lambda, double tau);
/* Integrator settings shared by all replicas (hmc.hpp:21-24). */
int ghx_set_integrator(ghx_ctx* ctx, int nmd, int integrator, int reunit_each_traj);
/* Multi-timescale integrator (the nested scheme of theory_notes.md:129-134;
 * the reference implements it only for U(1), u1.hpp:227-320): n_outer outer
 * steps of eps = tau/n_outer carrying the hopping forces (matter back-reaction
 * and the D(U) phi hops of the scalar force), leapfrog or Omelyan as chosen by
 * ghx_set_integrator, each inner half/whole step made of m_inner leapfrog
 * steps that carry the gauge force and the on-site scalar force. n_outer = 0
 * restores the single-scale integrator. */
int ghx_set_nested(ghx_ctx* ctx, int n_outer, int m_inner);
/* Integrator of the inner scale of ghx_set_nested: GHX_LEAPFROG (default) or
 * GHX_OMELYAN, the minimum-norm step the reference's U(1) nested scheme uses
 * on both scales (inner_fast, u1.hpp:300-317). m_inner Omelyan steps cost
 * 2*m_inner+1 fused inner launches, m_inner leapfrog steps m_inner+1. */
int ghx_set_nested_inner(ghx_ctx* ctx, int integrator);
/* Frozen-length scalar |phi_x| = 1 (GaugeHiggsHMC::frozen_phi,
 * gauge_higgs_hmc.hpp:38, 61-103, 133, 144, 172): momenta are projected onto
 * the tangent space at refresh and after every kick, and the phi drift is the
 * exact great-circle flow. ghx_normalize_phi is normalize_phi() (:74-84).
 * Not available with the large-irrep kernels (GHX_E_UNSUPPORTED). */
int ghx_set_frozen_phi(ghx_ctx* ctx, int on);
int ghx_normalize_phi(ghx_ctx* ctx);
/* Multi-invariant quartic potential (MultiInvariantPotential,
 * scalar_invariants.hpp:20-208, plugged in as GaugeHiggsHMC::potential):
 * V = -mu2 phi^dag phi + sum_c f_c ||P_c(phi phi^dag)||^2 over the channels c of
 * R x Rbar (eigenspaces of M -> sum_a [T^a_R,[T^a_R,M]]). ghx_casimir_channels
 * lists the channel Casimirs (host only; C2 may be NULL to get the count);
 * ghx_set_multi_potential needs one f_c per channel, f == NULL restores the
 * quartic potential. Not available with the large-irrep kernels. */
int ghx_casimir_channels(int N, const char* spec, double* C2, int* n);
int ghx_set_multi_potential(ghx_ctx* ctx, const double* f, int nf, double mu2);
/* RNG seed of one replica (Rng ctor, rng.hpp:21). */
int ghx_set_seed(ghx_ctx* ctx, int replica, uint64_t seed);

/* Start configuration of one replica: links cold or reference hot start with
 * sigma_U (fields.hpp:21-34); phi cold(c = phi_sigma) or Gaussian from
 * phi_stream with phi_sigma (scalar_field.hpp:33-51). */
int ghx_init_fields(ghx_ctx* ctx, int replica, int link_mode, double sigma_U, int phi_mode, uint64_t phi_stream,
                    double phi_sigma);

/* Trajectory/accept counters (hmc.hpp:26, gauge_higgs_hmc.hpp:42-43). */
int ghx_set_counters(ghx_ctx* ctx, int replica, uint64_t traj_count, uint64_t accept_count);
int ghx_get_counters(const ghx_ctx* ctx, int replica, uint64_t* traj_count, uint64_t* accept_count, double* last_dH);

/* Momentum heatbath of every replica with its current traj_count
 * (gauge_higgs_hmc.hpp:167-173, hmc.hpp:74-75). */
int ghx_refresh_momenta(ghx_ctx* ctx);
/* P -= eps F, pi += eps F_phi (gauge_higgs_hmc.hpp:112-134, hmc.hpp:35-42). */
int ghx_kick(ghx_ctx* ctx, double eps);
/* U <- exp(i eps P) U, phi += eps pi (gauge_higgs_hmc.hpp:136-147, hmc.hpp:44-52). */
int ghx_drift(ghx_ctx* ctx, double eps);
/* One MD trajectory with fused kick+drift launches (gauge_higgs_hmc.hpp:149-165). */
int ghx_md_evolve(ghx_ctx* ctx);
/* H parts per replica: [KE_P, KE_pi, S_g, S_H, H] x n_replicas (gauge_higgs_hmc.hpp:105-110). */
int ghx_hamiltonian(ghx_ctx* ctx, double* H_parts);
/* One full HMC trajectory of every replica with the Metropolis test on the GPU
 * (gauge_higgs_hmc.hpp:175-192, hmc.hpp:73-91). accepted/dH: n_replicas each, may be NULL. */
int ghx_trajectory(ghx_ctx* ctx, int* accepted, double* dH);
/* n trajectories; when use_graph, the first is captured into a hipGraph and replayed. */
int ghx_trajectories(ghx_ctx* ctx, int n, int use_graph);
/* Observables per replica: [plaquette, S_g, S_H, L_link, L_phi, Polyakov] x n_replicas
 * (gauge_wilson.hpp:51-65, scalar_higgs.hpp:130-154, observables.hpp:18-34). */
int ghx_observables(ghx_ctx* ctx, double* out);
/* Wilson loop traced in a probe irrep, per replica (observables.hpp:36-81, wilson_loop_fund /
 * wilson_loop_rep): W[r] = < (1/d_R') Re Tr D^(R')(U_loop) > over sites and planes mu < nu of the
 * Ra x Rb rectangle (Ra links along mu, Rb along nu). probe: "fund" (wilson_loop_fund), "adj" or
 * Young rows "r1[,r2][:real]"; an invalid spec gives GHX_E_REP with the GeneralRep message.
 * W has n_replicas entries. */
int ghx_wilson_loop(ghx_ctx* ctx, const char* probe, int Ra, int Rb, double* W);
/* Polyakov loop in direction tdir traced in a probe irrep, per replica (observables.hpp:83-101,
 * polyakov_loop_rep): < (1/d_R') Re Tr D^(R')(prod_t U_tdir) > over the sites with x_tdir = 0. */
int ghx_polyakov_loop_rep(ghx_ctx* ctx, const char* probe, int tdir, double* P);
/* Analysis of one observable's Monte Carlo series x[0..n) (host only, no context):
 * out[10] = {mean, population variance, Binder cumulant U4, block-jackknife error with
 * 16 bins (observables.hpp:104-145 Stats), tau_int with Sokal's window M >= 6 tau_int(M)
 * and lags capped at 2000, the window W, N_eff = n / (2 tau_int), naive error sqrt(Var/n),
 * autocorrelation-aware error sqrt(2 tau_int Var / n), blocking error with block size
 * `block`} (autocorr.hpp:55-212), with the reference's values for empty, length-1 and
 * constant series. */
int ghx_series_analysis(const double* x, long n, int block, double* out);
/* Normalised autocorrelation rho(t) = C(t)/C(0), t = 0..min(maxlag, n-1) (every lag when
 * maxlag < 0), biased 1/n autocovariance (autocorr.hpp:93-123); *nlag = entries written. */
int ghx_autocorr_function(const double* x, long n, int maxlag, double* rho, long* nlag);
/* max ||U^dag U - 1|| of one replica (fields.hpp:38-42). */
int ghx_max_unitarity(ghx_ctx* ctx, int replica, double* out);
/* Reunitarize all links (fields.hpp:36, sun.hpp:49-72). */
int ghx_reunitarize(ghx_ctx* ctx);

/* Copy fields of one replica to / from host AoS buffers (any pointer may be NULL). */
int ghx_get_fields(ghx_ctx* ctx, int replica, double* U, double* P, double* phi, double* pi);
int ghx_set_fields(ghx_ctx* ctx, int replica, const double* U, const double* P, const double* phi, const double* pi);
/* Link force F (gauge + matter, add_gauge_force + add_matter_link_force) and
 * scalar force F_phi (scalar_force) of one replica for the current fields
 * (gauge_wilson.hpp:71-83, scalar_higgs.hpp:67-128). */
int ghx_forces(ghx_ctx* ctx, int replica, double* F_link, double* F_phi);

/* Replica swap proposal between replicas i and j (same extents): exchanges U
 * and phi with probability min(1, exp(-Delta)) where Delta is the change of
 * the summed actions of both replicas (replica_tempering.hpp:171-194 for the
 * beta ladder; the kappa ladder uses Delta = (kappa_i-kappa_j)(B_i-B_j),
 * B = sum 2 Re phi^dag D phi'). u is the caller's uniform in (0,1). */
int ghx_replica_swap(ghx_ctx* ctx, int i, int j, double u, int* accepted, double* delta);

/* Replica exchange ladders (replica_tempering.hpp:78-205 generalised to beta
 * and kappa ladders): each ladder is an ordered list of distinct replicas with
 * equal extents and lambda; adjacent members are swap partners. A swap step
 * runs the even pass over all ladders then the odd pass, with the coin
 * uniform(swap_seed, key(0x57A9, ++counter)) (replica_tempering.hpp:171-194).
 * With tempering enabled, ghx_trajectories() runs one swap step after every
 * trajectory (inside the captured hipGraph). */
int ghx_set_ladders(ghx_ctx* ctx, int n_ladders, const int* lengths, const int* members, uint64_t swap_seed);
int ghx_set_tempering(ghx_ctx* ctx, int enabled);
int ghx_swap_step(ghx_ctx* ctx);
/* Pair list (even pass then odd pass, 2 ints per pair), cumulative attempts /
 * accepts per pair and the flags of the last swap step. Any pointer may be NULL. */
int ghx_swap_stats(ghx_ctx* ctx, int* npairs, int* pairs, uint64_t* attempts, uint64_t* accepts, int* last_flags);

/* Launch-shape tuning (no counterpart in the reference, which runs OpenMP
 * loops): GHX_TUNE_STREAM_SEG = number of slices of the last direction one
 * workgroup walks in the MD-step kernel (0 = one site per thread);
 * GHX_TUNE_BLOCK_ORDER = order in which the one-site-per-thread kernel walks the
 * blocks of a single lattice (0 storage order; 1 each XCD's share of slices
 * walked with the last direction fastest; 2 as 1 with direction D-2 next;
 * 3-6 a 4D tile of 128 sites per workgroup, 16x4x2x1, 8x4x2x2, 16x2x2x2 or
 * 32x2x2x1, for D = 4 lattices the tile divides; get_tuning reports 0 where the
 * order does not apply); GHX_TUNE_LINK_LANES = 1
 * runs the MD step with one lane per link (four per site) instead of one
 * thread per site, for lattices too small to fill the GPU; GHX_TUNE_SYM_ENGINE
 * = 1 (default where it applies) applies a large tensor irrep through
 * symmetric powers of the link, D(U) = Yc^T (Sym^k1(U) x Sym^k2(U)) Yc / c, instead of
 * slot by slot on the N^n tensor (get_tuning reports whether it is in use);
 * GHX_TUNE_SYM_PAIRS selects the engine's link kernel where SU(3) has a first
 * Young row <= 6: 2 (default) two links per wave with each lane's column of
 * Sym^k(U) in registers, the forward hop through Sym^k(U^T) (FP32 with a
 * second row <= 1; elsewhere it runs as 1), 1 two links per wave with
 * Sym^k(U) built a column per lane into LDS, 0 one link per wave with the
 * per-entry LDS program (get_tuning reports the kernel in use, 0 where neither
 * paired kernel applies); GHX_TUNE_X_LANES = 1 (default where it
 * applies: SU(3) in FP32 MD, L0 a power of two dividing 64, volumes and
 * cross-sections multiples of 64 sites) lets the one-site-per-thread and
 * streaming MD steps hand each site's lower staple in plane (0, mu) to its
 * x + 0 neighbour lane instead of forming it there (same physics, another
 * rounding; not with the 4D-tile walks or link lanes; get_tuning reports
 * whether the MD step uses it); GHX_TUNE_FUSE_INNER = 1 (default) runs the
 * inner-scale step of a large irrep under the multi-timescale integrator (the
 * scalar's on-site update and the links' gauge update, which touch disjoint
 * fields) as one launch where the links run link-parallel in D = 4, with the
 * arithmetic of the two separate launches (get_tuning reports whether it
 * applies); GHX_TUNE_SPLIT_INNER = 1 (default) runs a large irrep's whole
 * inner-scale block (the 2 m_inner + 1 or m_inner + 1 sub-steps of one
 * half step) as one launch for the scalar's on-site sub-steps and one
 * gauge-only launch per sub-step, where the two updates are independent: no
 * surrogate force in use, no plugged-in potential (same fields as the
 * interleaved launches; get_tuning reports whether it applies);
 * GHX_TUNE_STORE_POLICY = 0 (default) streams the MD step's stores past the
 * caches, 1 stores them with the default cache policy (faster where the next
 * step finds them in cache); GHX_TUNE_LINK_SPLIT = 1 (default) lets the
 * link-parallel step of a kernel without scalar components (pure gauge, or a
 * large irrep's gauge part) give each link D - 1 lanes, one per plane of its
 * staples (get_tuning reports whether it applies; another rounding of the
 * staple sum). The autotuner times every candidate launch shape
 * with both store policies (and the link-lane shape with and without the
 * plane split) for `steps` MD steps on the
 * live fields, restores the fields, and keeps the fastest; for a large irrep
 * whose fastest shape is the link-lane one, the plane split is then decided
 * on 8 x `steps` gauge-only launches of that step with and without it (the
 * launches a nested chain's inner scale makes), twice each. */
enum { GHX_TUNE_STREAM_SEG = 0, GHX_TUNE_BLOCK_ORDER = 1, GHX_TUNE_LINK_LANES = 2, GHX_TUNE_SYM_ENGINE = 3, GHX_TUNE_SYM_PAIRS = 4,
       GHX_TUNE_X_LANES = 5, GHX_TUNE_FUSE_INNER = 6, GHX_TUNE_SPLIT_INNER = 7, GHX_TUNE_STORE_POLICY = 8,
       GHX_TUNE_LINK_SPLIT = 9 };
int ghx_set_tuning(ghx_ctx* ctx, int key, int value);
int ghx_get_tuning(const ghx_ctx* ctx, int key, int* value);
int ghx_autotune(ghx_ctx* ctx, int steps, int* best_seg, double* best_ms_per_step);

/* Learned surrogate of the matter back-reaction on the links (no counterpart
 * in the reference): an MLP 16 -> 64 -> 64 -> 4 on BF16 MFMA with FP32
 * accumulation maps gauge-invariant link features to the coefficients of four
 * gauge-covariant su(N) vectors. With the surrogate enabled, the nested
 * integrator adds it to the inner (cheap) kicks and subtracts it in the outer
 * kicks, so the total force is exact on average and the MD stays reversible
 * and volume preserving; acceptance uses the exact FP64 H. */
int ghx_surrogate_init(ghx_ctx* ctx, uint64_t seed);
int ghx_surrogate_enable(ghx_ctx* ctx, int on);
int ghx_surrogate_nparams(void);
/* FP32 master parameters (get = 1 copies out, 0 copies in), ghx_surrogate_nparams() floats. */
int ghx_surrogate_params(ghx_ctx* ctx, int get, float* params);
/* Loss gradient on the current fields without updating (for checks). */
int ghx_surrogate_gradient(ghx_ctx* ctx, float* grad, double* loss);
/* `iters` Adam steps on the links of the current configuration (exact matter
 * force as the target); loss before the first and after the last step. */
int ghx_surrogate_train(ghx_ctx* ctx, int iters, double lr, double* loss_first, double* loss_last);
/* Features (rows*16), bases (rows*4*(N*N-1)), surrogate S and exact matter
 * force F (rows*(N*N-1)) of one replica, rows in the reference link order. */
int ghx_surrogate_snapshot(ghx_ctx* ctx, int replica, float* X, float* B, double* S, double* F);

/* Checkpoint / resume of every replica (roadmap item, README.md:126-127): links
 * and phi in the reference AoS layout (FP64), couplings, seeds, trajectory and
 * accept counters, last dH, the swap-coin counter, the per-pair swap attempts
 * and accepts, the frozen flag and the integrator settings, in a versioned
 * binary file. Loading checks D, N, the replica count, extents, d, the frozen
 * flag, the integrator settings and the number of swap pairs against the
 * context and returns GHX_E_INVALID on any mismatch. */
int ghx_save_checkpoint(ghx_ctx* ctx, const char* path);
int ghx_load_checkpoint(ghx_ctx* ctx, const char* path);

int ghx_sync(ghx_ctx* ctx);
/* Time `steps` fused MD steps (kick+drift launches) with HIP events on the
 * context's stream: total ms and the number of launches timed. */
int ghx_time_md_steps(ghx_ctx* ctx, int steps, double* ms_total);
/* Host-only representation query (no device needed): d, N-ality, real flag,
 * tensor dimension, and optionally W (d*dimT) and T^a_R ((N*N-1)*d*d complex);
 * the same construction ghx_set_rep uses (rep_general.hpp:54-84, 338-411). */
int ghx_rep_query(int N, const char* spec, int* d, int* nality, int* real_rep, long* dimT, double* W, double* T);
/* Total sites over all replicas. */
int64_t ghx_volume(const ghx_ctx* ctx);
/* The device's Gaussian stream: out[i] = Rng(seed).gauss(key0 + i) for
 * i < n (Rng::gauss, rng.hpp:42-46), drawn on the current HIP device by the
 * code the momentum refresh and the hot starts use, copied to host memory. */
int ghx_rng_gauss(uint64_t seed, uint64_t key0, int64_t n, double* out);
/* ghx_rng_gauss without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc); no counterpart in the reference. The draws are staged in
 * a file-scope device array of 2^20 doubles (8 MiB per GPU, held until the
 * process ends): n above 2^20 gives GHX_E_INVALID. Calls that overlap on one
 * GPU share that array and race. */
int ghx_rng_gauss_noalloc(uint64_t seed, uint64_t key0, int64_t n, double* out);

#ifdef __cplusplus
}
#endif
#endif /* GHX_H */
