/*
 * pycwb_amd -- C ABI of the MI355X-native coherent sky-scan likelihood and WDM
 * forward transform for PycWB.
 *
 * Every pointer argument is a device pointer (HBM) unless stated otherwise;
 * `stream` is a hipStream_t passed as void*.  Calls enqueue work on `stream` and
 * return without synchronising.  The only state kept between calls is the
 * text of the last error, which is per thread; concurrent calls on different
 * streams with different buffers are safe.
 * Return value: PYCWB_OK, or PYCWB_ERR_ARG (shapes / pointers rejected before
 * any launch), or PYCWB_ERR_HIP (a HIP runtime call failed).
 *
 * Array layouts follow the reference's Python call sites:
 *   FP, FX  (n_sky, n_ifo) float32  -- setup["FP_t"], setup["FX_t"]
 *   ml      (n_ifo, n_sky) int32    -- setup["ml"]
 *   rms     (n_pix, n_ifo) float32  -- noise_weights.T
 *   td00/90 (n_delay, n_ifo, n_pix) float32, n_delay = 2K+1 odd
 *   maps    (11, n_sky) float32, in the order of the reference's returned tuple:
 *           nAntennaPrior, nAlignment, nLikelihood, nNullEnergy, nCorrEnergy,
 *           nCorrelation, nSkyStat, nDisbalance, nNetIndex, nEllipticity,
 *           nPolarisation
 * Preconditions the host checks (the kernels only clamp for memory safety):
 *   0 <= sky_valid_indices[k] < n_sky, |ml| <= K, 1 <= n_ifo <= 8.
 */
#ifndef PYCWB_AMD_H
#define PYCWB_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYCWB_OK 0
#define PYCWB_ERR_ARG 1
#define PYCWB_ERR_HIP 2

/* flags of the regulator / scan entry points.  All are exact: results are
 * bit-identical with and without them, given finite noise weights and sky
 * tables (checked by the caller).
 *   PYCWB_SKIP_MASKED_PIXELS: a pixel whose energy at a direction's delays is
 *     below the network threshold adds nothing to any sum, so its evaluation
 *     stops after the energy test;
 *   PYCWB_REGULATOR_EARLY_EXIT: a direction's DPF sweep stops once its partial
 *     quality already exceeds gamma with every remaining term >= 0;
 *   PYCWB_TILE_ENERGY_BOUND (with PYCWB_SKIP_MASKED_PIXELS): each wavefront of
 *     64 directions bounds every pixel's energy over the taps its directions
 *     read and skips pixels whose bound is below the threshold for all of
 *     them, without loading their amplitudes per direction. */
#define PYCWB_SKIP_MASKED_PIXELS 1
#define PYCWB_REGULATOR_EARLY_EXIT 2
#define PYCWB_TILE_ENERGY_BOUND 4
/*   PYCWB_REGULATOR_BOUNDS (with PYCWB_REGULATOR_EARLY_EXIT): per-pixel bounds
 *     on the DPF quality terms, from the cluster's noise-weight range and the
 *     direction's antenna patterns, decide a direction's count before or
 *     during its sweep whenever the final sum is bracketed on one side of
 *     the threshold (margins cover float32 and float64 rounding). */
#define PYCWB_REGULATOR_BOUNDS 8

const char* pycwb_version(void);

/* Text of the last PYCWB_ERR_HIP raised on the calling thread (HIP error name
 * and the failing call). */
const char* pycwb_last_error(void);

/* Bytes of device workspace needed by pycwb_td_relayout / pycwb_sky_scan_plain
 * for one cluster (256 bytes of counters + the packed amplitude table). */
size_t pycwb_sky_scan_workspace_bytes(int n_ifo, int n_pix, int n_delay);

/* Packs td00/td90 (reference layout) into the workspace's (n_pix, n_ifo,
 * n_delay) {00,90} pair table read by the scan.  Replaces the per-cluster
 * transposes at pycwb/modules/likelihoodWP/likelihood.py:299-305. */
int pycwb_td_relayout(int n_ifo, int n_pix, int n_delay, const float* td00, const float* td90,
                      void* workspace, size_t workspace_bytes, void* stream);

/* DPF regulator pass.  Replaces calculate_dpf
 * (pycwb/modules/likelihoodWP/dpf.py:7-34; called at likelihood.py:310).
 * Writes REG[0] = reg0 (delta*sqrt(2) as float32), REG[1] = (float)regulator,
 * REG[2] = 0, and the unrounded float64 regulator to *reg_f64 when non-NULL.
 * workspace needs >= 256 bytes. */
int pycwb_dpf_regulator(int n_ifo, int n_pix, int64_t n_sky, const float* FP, const float* FX,
                        const float* rms, double gamma_regulator,
                        double network_energy_threshold, float reg0,
                        const int64_t* sky_valid_indices, int64_t n_valid, int flags,
                        float* REG, double* reg_f64, void* workspace, size_t workspace_bytes,
                        void* stream);

/* Diagnostic without a counterpart in the reference: the term bounds the
 * regulator passes decide with (PYCWB_REGULATOR_BOUNDS).  From the cluster's
 * noise weights rms ((n_pix, n_ifo) float32) writes its weight statistics
 * (24 doubles: min and max of R2 = sum_i r_i^2, then the per-detector minima
 * and maxima of r_i^2 / R2 in slots 2.. and 10..) to weight_stats, and for
 * each of n_dirs directions (rows of FP, FX: (n_dirs, n_ifo) float32) the
 * lower bound t_lb and upper bound t_ub (< 0: none) on every pixel's term
 * fx / (ni + 1e-9).  For tests that hold the kernel's bounds against the
 * float32 terms. */
int pycwb_regulator_term_bounds(int n_ifo, int n_pix, const float* rms, const float* FP,
                                const float* FX, int64_t n_dirs, double* weight_stats,
                                double* t_lb, double* t_ub, void* stream);

/* Diagnostic without a counterpart in the reference: the pixel classes of
 * one cluster (pixel_class, n_pix bytes; class_stats, 16 x 24 doubles: per
 * class the weight statistics in the layout of weight_stats, the pixel count
 * in slot 18 and, in class 0, the number of classes in slot 19; class_rem,
 * pycwb_class_rem_entries(n_pix) uint16: row q, class g = pixels j >= 16 q
 * of class g) and, for each of n_dirs directions, every class's term bounds
 * (t_lb, t_ub: n_dirs x 16; 0 and -1 for unused classes), as the batched
 * regulator's sweep uses them. */
int pycwb_regulator_class_bounds(int n_ifo, int n_pix, const float* rms, const float* FP,
                                 const float* FX, int64_t n_dirs, double* weight_stats,
                                 uint8_t* pixel_class, double* class_stats, uint16_t* class_rem,
                                 double* t_lb, double* t_ub, void* stream);

/* Sky scan of one cluster.  Replaces scan_sky_for_best_fit
 * (pycwb/modules/likelihoodWP/sky_scan.py:17-192; called at likelihood.py:319).
 * Reads the amplitude pairs from td_packed ((n_pix, n_ifo, n_delay) float
 * pairs) or, when td_packed is NULL, from the workspace filled by
 * pycwb_td_relayout (workspace >= 256 bytes suffices with td_packed).
 * sky_order (may be NULL) is a permutation of valid-list positions: the order
 * in which lanes take directions (directions sharing delay taps side by side
 * keep PYCWB_SKIP_MASKED_PIXELS branches coherent); results are indexed by
 * sky pixel and tie-broken by valid-list position either way.  maps may be NULL
 * (argmax only).  Writes *l_max (int64) and *sky_stat_max (float64) on the
 * device.  A direction with no pixel above the energy threshold is treated as
 * gated (the reference divides 0/0 there). */
int pycwb_sky_scan_plain(int n_ifo, int n_pix, int64_t n_sky, int n_delay, const float* FP,
                         const float* FX, const int32_t* ml, const float* rms,
                         const float* td_packed, const float* REG, double netCC,
                         double network_energy_threshold, const int64_t* sky_valid_indices,
                         int64_t n_valid, const int32_t* sky_order, int flags, float* maps,
                         int64_t* l_max, double* sky_stat_max, void* workspace,
                         size_t workspace_bytes, void* stream);

/* Batched regulator + scan + argmax over many clusters, as one hipGraph of
 * persistent work-queue kernels.  Replaces the per-cluster loop of
 * evaluate_fragment_clusters / likelihood_wrapper
 * (pycwb/modules/likelihoodWP/likelihood.py:58-135) up to the scan.
 * Clusters are listed in `order`, grouped by detector count: group g covers
 * order[group_first[g] .. group_first[g]+group_count[g]) whose clusters all
 * have group_nifo[g] detectors.  Per cluster c:
 *   rms at rms + rms_off[c]  (n_pix[c], n) float32
 *   td  at td + 2*td_off[c]  (n_pix[c], n, 2K[c]+1) {00,90} float pairs
 *   network net[c]: FP/FX at FP/FX + net_off[net] (n_sky, n), ml at
 *   ml + net_off[net] (n, n_sky), energy threshold net_En[net].
 * Scratch: tickets (16*n_groups uint32: per group and phase, one work queue
 * per XCD), counts, best (per cluster).
 * Outputs per cluster: REG1 (float32), l_max (int64), sky_stat_max (float64),
 * reg_f64 (float64, may be NULL).  n_sky < 2^31 - 1. */
typedef struct pycwb_batch_desc {
  int n_clusters;
  int64_t n_sky;
  const int64_t* sky_valid_indices;
  int64_t n_valid;
  const int32_t* order;
  int n_groups;
  int group_first[8];
  int group_count[8];
  int group_nifo[8];
  const int32_t* n_pix;
  const int32_t* net;
  const int32_t* K;
  const int64_t* rms_off;
  const int64_t* td_off;
  const float* rms;
  const float* td;
  const int32_t* net_nifo;
  const int64_t* net_off;
  const double* net_En;
  const float* FP;
  const float* FX;
  const int32_t* ml;
  const int32_t* sky_order; /* NULL or (n_networks, n_valid) valid-list positions */
  int flags;                /* PYCWB_SKIP_MASKED_PIXELS | PYCWB_REGULATOR_EARLY_EXIT */
  double gamma_regulator;
  float reg0;
  double netCC;
  uint32_t* tickets;
  uint32_t* counts;
  unsigned long long* best;
  float* REG1;
  int64_t* l_max;
  double* sky_stat_max;
  double* reg_f64;
  int block;      /* sky directions per workgroup: 64, 128 or 256 */
  int grid;       /* persistent workgroups */
  int use_graph;  /* capture the launches in a hipGraph */
  double* weight_stats; /* scratch, 24 doubles per cluster, or NULL (no regulator bounds) */
  /* Pixel classes for the regulator sweep (all five, with weight_stats and
   * PYCWB_REGULATOR_EARLY_EXIT | PYCWB_REGULATOR_BOUNDS, or NULL: none).
   * Scratch: pixel_class one byte per noise weight (sum of n_pix[c] times
   * the cluster's detector count; cluster c's n_pix[c] class bytes at
   * rms_off[c]), class_stats 384 doubles per cluster, class_rem
   * pycwb_class_rem_entries(n_pix[c]) uint16 per cluster from
   * class_rem_off[c] (a multiple of 16; the buffer 32-byte aligned). */
  uint8_t* pixel_class;
  double* class_stats;
  uint16_t* class_rem;
  const int64_t* class_rem_off;
  /* class-bound scratch of at least pycwb_class_bound_scratch_bytes(block,
   * grid) bytes (class_bounds_bytes), 16-byte aligned */
  uint32_t* class_bounds;
  size_t class_bounds_bytes;
} pycwb_batch_desc;

/* uint16 entries of class_rem one cluster of n_pix pixels needs. */
int64_t pycwb_class_rem_entries(int n_pix);

/* Bytes of class-bound scratch a batch of this launch shape needs. */
size_t pycwb_class_bound_scratch_bytes(int block, int grid);

typedef struct pycwb_batch pycwb_batch;

int pycwb_batch_create(const pycwb_batch_desc* desc, pycwb_batch** out);
int pycwb_batch_launch(pycwb_batch* batch, void* stream);
/* One stage of the batch outside the graph (0 init, 1 regulator, 2 scan,
 * 3 finish), for per-stage timing; stages must run in order.  Stage 4, run
 * right after stage 0 instead of 1-3, writes into the per-cluster counts the
 * number of (direction, pixel) pairs above the energy threshold (the pairs
 * the scan evaluates in full); stage 5, likewise after stage 0 and once a
 * launch has finished, the number of pairs whose energy the scan evaluates
 * (the others are ruled out by its tile energy bound).  Both are measurement
 * aids. */
int pycwb_batch_launch_stage(pycwb_batch* batch, int stage, void* stream);
int pycwb_batch_destroy(pycwb_batch* batch);

/* The batched regulator + scan + argmax of pycwb_batch_create/launch as one
 * call that allocates nothing: no hipMalloc, no hipGraph, no host heap.  No
 * counterpart in the reference; for callers that cannot allocate (code under
 * stream capture, serving loops, harnesses that forbid allocation).  The
 * kernels and results are those of pycwb_batch_launch on the same description;
 * the description's tickets, counts, best, weight_stats and pixel-class
 * buffers are ignored (may be NULL; the regulator then bounds its sweep by
 * the cluster-wide bounds only, with the same counts): the work queues,
 * regulator counts, argmax keys and noise-weight statistics live in a device
 * array of the library (about 3.3 MB per GPU,
 * held until the process ends), cleared by a fill kernel at the start of
 * every call.  At most pycwb_batch_noalloc_max_clusters() (16,384) clusters;
 * more return PYCWB_ERR_ARG with the reason in pycwb_last_error().  The
 * kernels are enqueued on `stream` directly (use_graph is ignored).  Calls
 * that overlap on one GPU share that array and race: serialise them. */
int pycwb_batch_run_noalloc(const pycwb_batch_desc* desc, void* stream);
size_t pycwb_batch_noalloc_max_clusters(void);

/* Measurement aid, no counterpart in the reference: a kernel of known VALU
 * instruction count (blocks x 4 wavefronts, each 64 * iters float32 FMAs in
 * eight independent chains, no memory traffic) whose run time gives the
 * device's VALU issue ceiling in wave-instructions per second.  out needs
 * `blocks` floats (never written in practice). */
int pycwb_valu_probe(float* out, int blocks, int iters, void* stream);

/* Best-sky statistics of a batch of clusters at their selected directions,
 * one workgroup per cluster.  Replaces compute_statistics_at_sky_position
 * (pycwb/modules/likelihoodWP/sky_statistics.py:26-290; called at
 * likelihood.py:359) with its packet operations (packet_ops.py).
 * Cluster c uses the batch layout of pycwb_batch_desc (rms, td, sky tables),
 * its direction l_max[c] and regulator REG1[c] (REG[0] = reg0), and its
 * cross-talk CSR: lookup + 2*lk_off[c] ((n_pix, 2) int32 [start, end) rows into
 * xt + 8*xt_off[c], whose (m, 8) float32 rows hold the neighbour's local pixel
 * index in column 0 and the coefficients in columns 4..7).
 * Scratch per cluster: s32 + s32_off[c] with pycwb_sky_stats_scratch32(n, P)
 * floats and s64 + s64_off[c] with pycwb_sky_stats_scratch64(n, P) doubles;
 * the reconstructed packets stay there (layout in pycwb_amd/skystats.py).
 * Outputs: scal (16 doubles per cluster: Gn, Ec, Dc, Rc, Eh, Es, Np, Em, Lm,
 * norm, cc, rho, xrho, Lo, Eo, N_pix_effective) and S_snr (8 per cluster).
 * Launches n_clusters workgroups; block b handles cluster which[b] (or b when
 * which is NULL). */
typedef struct pycwb_stats_desc {
  int n_clusters;
  const int32_t* which;
  int64_t n_sky;
  const int32_t* n_pix;
  const int32_t* nifo;
  const int32_t* net;
  const int32_t* K;
  const int64_t* rms_off;
  const int64_t* td_off;
  const float* rms;
  const float* td;
  const int64_t* net_off;
  const double* net_En;
  const float* FP;
  const float* FX;
  const int32_t* ml;
  const int64_t* l_max;
  const float* REG1;
  float reg0;
  const int64_t* lk_off;
  const int32_t* lookup;
  const int64_t* xt_off;
  const float* xt;
  int xgb_rho_mode;
  float* s32;
  const int64_t* s32_off;
  double* s64;
  const int64_t* s64_off;
  double* scal;
  double* S_snr;
} pycwb_stats_desc;

size_t pycwb_sky_stats_scratch32(int n_ifo, int n_pix);
size_t pycwb_sky_stats_scratch64(int n_ifo, int n_pix);
int pycwb_sky_statistics(const pycwb_stats_desc* desc, void* stream);

/* Continuous-direction sky statistic and its gradients, float64, one
 * workgroup per cluster.  Extends the sky scan (pycwb/modules/likelihoodWP/
 * sky_scan.py:131-141, the statistic AA = L * Co maximised by the grid argmax)
 * to off-grid directions: delays tau_i = (R_i - R_ref).n(theta, phi) * rate / c
 * are linearly interpolated in the td tables (instead of rint at
 * pycwb/types/detector.py:1299-1303) and F+/Fx come from the response tensor
 * (detector.py:272-378) at psi = 0, gha = -phi; REG is fixed and the netCC
 * gate is not applied.  No counterpart exists in the reference; it serves
 * sky-location refinement off the HEALPix grid.
 *
 * theta = colatitude, phi = Earth-fixed longitude (radians), per cluster.
 * geo: per network 8 x 12 float64: for each detector the delay vector
 * (R_i - R_ref) * rate / c (3) then the response tensor row-major (9).
 * value[c] = AA, grad[2c..2c+1] = (dAA/dtheta, dAA/dphi).
 * grad_td (optional): float32 pairs in the td layout; on return holds
 * grad_scale[c] * dAA/d(amplitude) at the two taps each (pixel, detector)
 * reads and is left untouched elsewhere (callers zero it first).
 * A cluster whose network does not have n_ifo detectors gets NaN outputs. */
typedef struct pycwb_grad_desc {
  int n_clusters;            /* clusters evaluated by this call */
  const int32_t* clusters;   /* NULL: 0 .. n_clusters-1; else their indices */
  int n_ifo;                 /* detectors of every evaluated cluster (1..5) */
  const int32_t* n_pix;
  const int32_t* net;
  const int32_t* K;
  const int64_t* rms_off;
  const int64_t* td_off;     /* in float pairs */
  const float* rms;
  const float* td;
  const double* geo;
  const int32_t* net_nifo;
  const double* net_En;
  const double* REG1;        /* per cluster */
  double reg0;               /* delta * sqrt(2) */
  const double* theta;
  const double* phi;
  double* value;
  double* grad;
  float* grad_td;            /* NULL: angles only */
  const double* grad_scale;  /* NULL: 1 */
} pycwb_grad_desc;

int pycwb_sky_grad(const pycwb_grad_desc* desc, void* stream);

/* Time-delay amplitudes of n_pixels pixels of one detector's WDM map.
 * Replaces batch_get_td_vecs (pycwb/utils/td_vector_kernels.py:131-199,
 * called through TDBatchInputs.extract_td_vecs, types/td_batch_inputs.py:45-67).
 * padded00/90: (n_rows, M+1) float32 planes, the map's 00/90 phases zero-padded
 * by n_coeffs time bins on both sides (time_frequency_map.py:482-521);
 * T0, Tx: TD filters (WDM<T>::setTDFilter) stored tap-major, (2 n_coeffs+1,
 * 2J+1) float64: element [k][dT + J] is tap k - n_coeffs of delay dT (the
 * transpose of the reference's (2J+1, 2 n_coeffs+1) tables, so that threads of
 * consecutive delays read consecutive addresses);
 * trig: 4J (sin, cos) float64 pairs of q pi / (2J), q = 0..4J-1 (the phase
 * table of WDM<T>::setTDFilter, cwb-core/WDM.cc:688-699);
 * pixel_index[p] = n (M+1) + m; delays dT = -K..K.
 * out + p * row_stride receives, float32, either [a00(-K..K), a90(-K..K)]
 * (interleave = 0, the reference's rows) or (a00, a90) pairs per delay
 * (interleave = 1, the likelihood's packed layout).  Rows outside the planes
 * (whole-bin shifts at the map's ends) read as zero. */
int pycwb_td_vectors(const int32_t* pixel_index, int64_t n_pixels, const float* padded00,
                     const float* padded90, int64_t n_rows, int M, const double* T0,
                     const double* Tx, const double* trig, int n_coeffs, int J, int K,
                     float* out, int64_t row_stride, int interleave, void* stream);

/* Sub-network cut of a batch of clusters: the loudest pixels of each
 * supercluster (apply_subnet_cut, pycwb/modules/super_cluster_native/
 * utils.py:640-755, via _sub_net_cut_prepared_packets, sub_net_cut.py:206-267).
 * Cluster c uses the layout of pycwb_batch_desc: rms (n_pix[c], n) float32 at
 * rms + rms_off[c]; td (n_pix[c], n, 2K[c]+1) {00,90} float pairs at
 * td + 2*td_off[c]; network net[c]: FP/FX (n_sky, n) and ml (n, n_sky) at
 * offset net_off[net], with |ml| <= K[c]; thresholds net_En[net] =
 * float32(2 Acore^2 n) and net_Es[net] = float32(2 e2or).
 * A call evaluates the clusters which[0 .. n_which) (n_which <= 65535), all
 * with n_ifo detectors (1..8). */
typedef struct pycwb_subnet_desc {
  int n_which;
  const int32_t* which;
  int n_ifo;
  int64_t n_sky;
  const int32_t* n_pix;
  const int32_t* net;
  const int32_t* K;
  const int64_t* rms_off;
  const int64_t* td_off;     /* in float pairs */
  const float* rms;
  const float* td;
  const int64_t* net_off;
  const float* net_En;
  const float* net_Es;
  const float* FP;
  const float* FX;
  const int32_t* ml;
  double subcut;             /* < 0 disables the sub-network pre-cut */
  double* tiles;             /* scratch: n_which * pycwb_subnet_tiles(n_sky) * 8 doubles */
  double* sky;               /* 8 doubles per cluster (indexed by cluster) */
  /* pycwb_subnet_mra only: cross-talk CSR as in pycwb_stats_desc, scratch, output */
  const int64_t* lk_off;
  const int32_t* lookup;
  const int64_t* xt_off;
  const float* xt;
  float* scratch;            /* pycwb_subnet_mra_scratch(n, n_pix[c]) floats at scr_off[c] */
  const int64_t* scr_off;
  double* mra;               /* 8 doubles per cluster (indexed by cluster) */
  /* NULL, or per network a permutation of the n_sky directions ((n_networks,
   * n_sky) int32 at n_sky * net): lanes take directions in that order (group
   * directions with similar delays; results do not depend on it) */
  const int32_t* sky_order;
} pycwb_subnet_desc;

size_t pycwb_subnet_tiles(int64_t n_sky);
size_t pycwb_subnet_mra_scratch(int n_ifo, int n_pix);

/* Sky loop of the sub-network cut.  Replaces optimze_sky_loc_from_td
 * (pycwb/modules/super_cluster_native/sub_net_cut.py:404-517).  Writes
 * sky[8c .. 8c+7] = (l_max, stat, Em, Am, lm, Vm, suball, EE), the reference's
 * return tuple (zeros when no direction has a positive statistic). */
int pycwb_subnet_sky(const pycwb_subnet_desc* desc, void* stream);

/* Multi-resolution statistics at l_max = sky[8c].  Replaces
 * mra_statistics_from_td (sub_net_cut.py:659-724) with sse_MRA_ps (:746-825).
 * Cross-talk rows must not repeat a neighbour.  Writes mra[8c .. 8c+7] =
 * (submra, rHo, Eo, Lo, Ls, m, extracted components, pixels above En). */
int pycwb_subnet_mra(const pycwb_subnet_desc* desc, void* stream);

/* Cross-talk CSR of a batch of clusters from the WDM cross-talk catalog.
 * Replaces getXTalk_pixels_fast (pycwb/modules/xtalk/monster.py:240-440) as
 * called through XTalk.get_xtalk_pixels / get_xtalk_pixels_from_arrays
 * (pycwb/modules/xtalk/type.py:34-63).  Cluster c owns pixels
 * [pix_off[c], pix_off[c+1]) with nlayer = layers + 1 and time = WDM map index
 * (time bin * nlayer + frequency).  Catalog (load_catalog, monster.py:63-153):
 * layers (n_res) int32 > 0, table (n_res, n_res, max_layer+1, 2, 2) int32
 * [start, end) ranges into coeff (n_coeff, 5) float32.
 * Workspace: hash_keys (int64) and hash_vals (int32), n_slots entries, cluster
 * c owning [hash_off[c], hash_off[c+1]) -- a power of two >= 2 * its pixels;
 * counts (n_pixels int32), row_off (n_pixels + 1 int64), block_sums
 * (pycwb_xtalk_scan_blocks(n_pixels) int64), sort_keys (xt_capacity int64).
 * pycwb_xtalk_count fills row_off and xt_off (n_clusters + 1 int64: each
 * cluster's first entry; xt_off[n_clusters] = total entries); the host reads
 * the total, sizes xt, and pycwb_xtalk_fill writes xt (total, 8) float32 rows
 * in the reference's order and lookup (n_pixels, 2) int32 [start, end)
 * relative to the cluster's first entry -- each cluster's slices equal the
 * reference's (lookup, xtalk) pair. */
typedef struct pycwb_xtalk_desc {
  int n_clusters;
  int64_t n_pixels;
  int64_t n_slots;
  const int64_t* pix_off;
  const int32_t* nlayer;
  const int64_t* time;
  const int64_t* hash_off;
  int n_res;
  int max_layer;
  const int32_t* layers;
  const int32_t* table;
  const float* coeff;
  int64_t n_coeff;
  int64_t* hash_keys;
  int32_t* hash_vals;
  int32_t* counts;
  int64_t* row_off;
  int64_t* sort_keys;
  int64_t* block_sums;
  int32_t* lookup;
  int64_t* xt_off;
  float* xt;
  int64_t xt_capacity;
} pycwb_xtalk_desc;

size_t pycwb_xtalk_scan_blocks(int64_t n_pixels);
int pycwb_xtalk_count(const pycwb_xtalk_desc* desc, void* stream);
int pycwb_xtalk_fill(const pycwb_xtalk_desc* desc, void* stream);

/* WDM forward transform t2w of n_series float64 series (row stride ldx), all
 * with the same filter.  Replaces WDM<T>::t2w (cwb-core/WDM.cc:1385-1534) as
 * called through wdm_wavelet's t2w at
 * pycwb/modules/coherence_native/tf_batch_generation.py:111-191 and
 * pycwb/utils/td_vector_batch.py:72-80.
 * M = layers (power of two, 2M <= 2048), stride = time step MM (M for the
 * quadrature transform), filter = m_H taps (WDM<T>::getFilter),
 * twiddle = 2M complex float64 values exp(-2 pi i j / 2M), j = 0..2M-1.
 * quadrature != 0: out holds (n_series, N, M+1) complex pairs (map00, map90);
 * quadrature == 0: out holds the power map in the real parts.
 * N = ceil(n_samples / stride).  Requires n_samples > m_H. */
int pycwb_wdm_t2w(const double* x, int64_t n_samples, int64_t ldx, int n_series, int M,
                  int stride, const double* filter, int m_H, const double* twiddle,
                  int quadrature, double* out, void* stream);

/* pycwb_wdm_t2w (stride M) of n_shifts shifted copies of one series ts,
 * without the copies being written: copy s is ts shifted by shifts[s] as
 * pycwb_max_energy_shift defines it (k > 0: x[i] = ts[i + k] for i < L - k,
 * k < 0: x[i] = ts[i + k] for i >= -k, the other end keeping the original
 * samples).  The maps equal pycwb_wdm_t2w's of the copies.  Replaces the
 * shifted copies' transforms in the delay loop of time_delay_max_energy
 * (pycwb/modules/coherence_native/time_delay_numba.py:183-288, which shifts
 * the series and transforms each copy with wdm_wavelet's t2w).  Requires
 * m_H = 12 M + 1 (the WDM filters of precision 10 and their relatives);
 * PYCWB_ERR_ARG otherwise.  shifts: device int64, n_shifts entries. */
int pycwb_wdm_t2w_shifted(const double* ts, int64_t n_samples, const int64_t* shifts,
                          int n_shifts, int M, const double* filter, int m_H,
                          const double* twiddle, int quadrature, double* out, void* stream);

/* WDM inverse transforms w2t (quad = 0: from the 00 coefficients) and w2tQ
 * (quad = 1: from the 90 coefficients) of n_series maps.  Replaces
 * WDM<T>::w2t / w2tQ (cwb-core/WDM.cc:1536-1710) as called through
 * wdm_wavelet's w2t / w2tQ (pycwb/modules/data_conditioning/whitening.py:110-111,
 * PSD_correction.py:130-131, regression.py:801-806).
 * map: (n_series, N, M+1) complex pairs (00, 90) -- pycwb_wdm_t2w's output
 * layout; M a power of two, 2 <= M <= 1024; filter/twiddle as for t2w;
 * work: (n_series, N, 2M) float64 scratch; out: (n_series, n_out) float64,
 * the first n_out samples of the reconstruction. */
int pycwb_wdm_w2t(const double* map, int64_t N, int n_series, int M, const double* filter,
                  int m_H, const double* twiddle, int quad, double* work, double* out,
                  int64_t n_out, void* stream);

/* Network-pixel selection of n_lag lags at once.  Replaces
 * select_network_pixels (pycwb/modules/coherence_native/selection.py:144-243)
 * with _align_threshold_map_numba, _candidate_passes_support_numba and
 * _select_candidates_numba (coherence_native/kernels.py:107-137,231-368),
 * called once per lag by coherence_single_lag (pipeline.py:81-172).
 * energy: (n_ifo, n_freq, n_time) float64 per-detector pixel energies;
 * shift_bins: (n_lag, n_ifo) circular time shifts (>= 0) within the valid
 * range [edge_bins, n_time - edge_bins); veto: (n_time) int16 (1 keep) or
 * NULL; ib >= 1 / ie: first and end frequency rows; eo: the energy threshold.
 * Workspace: combined (n_lag, n_freq, n_time) float64, flags (same) uint8,
 * live (n_lag, n_time) uint8 (the shifted-veto live mask, an output too),
 * row_off (n_lag * n_freq + 1) int64, lag_off (n_lag + 1) int64.
 * pycwb_select_count fills them (lag_off[n_lag] = total pixels); the host
 * reads the total, sizes the outputs (capacity pixels) and pycwb_select_emit
 * writes, per lag from lag_off[l], the pixels in (frequency, time) order:
 * freq, time (int64), value (summed raw energy), det_energy (n_ifo each,
 * clipped at 0) and det_index (n_ifo each, time * n_freq + frequency). */
typedef struct pycwb_select_desc {
  int n_ifo;
  int n_freq;
  int64_t n_time;
  const double* energy;
  int n_lag;
  const int64_t* shift_bins;
  const int16_t* veto;
  int edge_bins;
  int ib;
  int ie;
  double eo;
  double* combined;
  uint8_t* flags;
  uint8_t* live;
  int64_t* row_off;
  int64_t* lag_off;
  int64_t capacity;
  int64_t* freq;
  int64_t* time;
  double* value;
  double* det_energy;
  int64_t* det_index;
} pycwb_select_desc;

int pycwb_select_count(const pycwb_select_desc* desc, void* stream);
int pycwb_select_emit(const pycwb_select_desc* desc, void* stream);

/* Time-delay max-energy maps (pattern path).  Replaces the delay loop of
 * time_delay_max_energy_numba (pycwb/modules/coherence_native/
 * time_delay_numba.py:183-288, 578-679; called through max_energy,
 * projection.py:63-124, by the coherence setup, setup.py:171-195).
 * pycwb_max_energy_shift: n_copies shifted copies (n_copies, L) of a series,
 * shift k > 0: left (x[i] = ts[i + k] for i < L - k), k < 0: right
 * (x[i] = ts[i + k] for i >= -k), the other end keeping the original samples.
 * pycwb_max_energy_packet: packet energy (_wdm_packet_energy_nb, :49-104) of
 * n_maps quadrature maps (n_maps, T, M1) (00, 90) pairs -- pycwb_wdm_t2w's
 * layout -- folded into the running pixel maximum cur (T, M1) float64
 * (init = 1: the first map initialises it).  jb/je/mL/mH/mean/p are
 * _compute_packet_energy_params (time_delay_packet.py:8-64).
 * pycwb_max_energy_finish zeroes layer 0 (and 1 when zero_two_layers). */
typedef struct pycwb_packet_desc {
  const double* maps;
  int n_maps;
  int64_t T;
  int M1;
  int64_t jb;
  int64_t je;
  int mL;
  int mH;
  double mean;
  int64_t p[9];
  int init;
  double* cur;
} pycwb_packet_desc;

int pycwb_max_energy_shift(const double* ts, int64_t L, const int64_t* shifts, int n_copies,
                           double* out, void* stream);
int pycwb_max_energy_packet(const pycwb_packet_desc* desc, void* stream);
int pycwb_max_energy_finish(double* cur, int64_t T, int M1, int zero_two_layers, void* stream);

/* ---- Connected-component clustering of the selected pixels ----
 * Replaces cluster_pixels (pycwb/modules/coherence_native/clustering.py:
 * 18-157) for every lag of a call: _label_components_grid (kernels.py:24-105)
 * and _subnet_subrho_batch_numba (kernels.py:194-229) with noise_rms = 1.
 * Pixels: lag_off (n_lag + 1) int64 ranges of lag-major pixel lists (the
 * layout pycwb_select_emit writes); freq, time int64 TF bins; energy (n_pix)
 * and det_energy (n_pix, n_ifo) float64; noise_rms (n_pix, n_ifo) or NULL
 * (all 1).  kf, kt >= 1 connectivity radii; dt = 1 / rate (1 when rate <= 0),
 * half_rate = rate / 2, n_sub = 2 iGamma^-1(n_ifo - 1, 0.314).
 * Outputs: label (n_pix) 1-based within the lag; lag_cluster_off (n_lag + 1)
 * cluster ranges per lag; cluster_off (n_pix + 1 capacity) and members
 * (n_pix): the pixels of each cluster in index order; per cluster (n_pix
 * capacity) c_energy, c_time, c_freq (numpy's summation order), c_subnet,
 * c_subrho.  work / work_bytes: device workspace of at least
 * pycwb_cluster_work_bytes (nine int32 and four float64 values per pixel and
 * hipCUB's temporary storage).  index_map: int32 (n_lag, n_freq, n_time) cells holding -1 on
 * entry (pycwb_cluster_map_init sets them once; every call leaves them -1
 * again, so one map serves every later call of that shape or smaller on the
 * same stream).  One call is stream-ordered and needs no host
 * synchronisation; the cluster count is lag_cluster_off[n_lag].  Limits:
 * n_lag <= 65535, n_pix < 2^31 - 1. */
typedef struct pycwb_cluster_desc {
  int n_lag;
  const int64_t* lag_off;
  int64_t n_pix;
  const int64_t* freq;
  const int64_t* time;
  const double* energy;
  const double* det_energy;
  const double* noise_rms;
  int n_ifo;
  int n_freq;
  int n_time;
  int kf;
  int kt;
  double dt;
  double half_rate;
  double n_sub;
  void* work;
  size_t work_bytes;
  int32_t* index_map;
  int64_t* label;
  int64_t* lag_cluster_off;
  int64_t* cluster_off;
  int64_t* members;
  double* c_energy;
  double* c_time;
  double* c_freq;
  double* c_subnet;
  double* c_subrho;
} pycwb_cluster_desc;

int pycwb_cluster_map_init(int32_t* index_map, int64_t cells, void* stream);
int pycwb_cluster_work_bytes(const pycwb_cluster_desc* desc, size_t* bytes);
int pycwb_cluster_pixels(const pycwb_cluster_desc* desc, void* stream);

/* ---- Pixel-energy threshold statistics ----
 * Replaces the array work of compute_threshold
 * (pycwb/modules/coherence_native/veto_threshold.py:98-222); the Gamma fits on
 * the resulting scalars stay on the host.  maps (n_ifo, n_freq, n_time)
 * float64 are summed over detectors in order and clipped to [0, hi]
 * (hi = n_ifo * 100) into values (count), in the reference's order:
 * time_major = 0: rows of frequency over time bins [begin, begin + count /
 * n_freq) (the edge-cropped sum raveled, :173-176); time_major = 1: the
 * time-major flat indices [begin, begin + count) (t * n_freq + f; the
 * w[nL:nR] region, :136-157).  counts[0], counts[1]: values above 1e-4 and
 * above 1e-3.  sorted (count) or NULL: the values ascending (np.sort,
 * :186-201).  chunk_sums (2 * ceil(count / 8192)) or NULL: the values above
 * 1e-3, in order, summed per 8192-value chunk in numpy's pairwise order, with
 * the sum of their logs beside each (np.mean(positive), np.mean(np.log(
 * positive)), :162-163, once the host adds the counts[1] / 8192 rounded up
 * chunk sums to 0.0 in order).  work: at least pycwb_threshold_work_bytes.
 * begin + count < 2^31 - 1. */
typedef struct pycwb_threshold_desc {
  const double* maps;
  int n_ifo;
  int n_freq;
  int64_t n_time;
  int time_major;
  int64_t begin;
  int64_t count;
  double hi;
  void* work;
  size_t work_bytes;
  double* values;
  uint64_t* counts;
  double* sorted;
  double* chunk_sums;
} pycwb_threshold_desc;

int pycwb_threshold_work_bytes(const pycwb_threshold_desc* desc, size_t* bytes);
int pycwb_threshold_stats(const pycwb_threshold_desc* desc, void* stream);

/* ---- Gamma-to-Gauss of a max-energy map ----
 * Replaces the array work of TimeFrequencyMap.Gamma2Gauss
 * (pycwb/types/time_frequency_map.py:143-246, called by
 * time_delay_max_energy_numba, coherence_native/time_delay_numba.py:664);
 * the Gamma fit on the resulting scalars stays on the host.  flat: the map
 * time-major (n = layers * time bins, flat[t * layers + f]), region
 * [nL, nR), window [nL, nR].  Phases, in order, each after the host has read
 * the previous one's results:
 *   1: count[0] = #(flat > 0.001); sorted = the window of flat, ascending
 *      (the median is sorted[split - nL]);
 *   2: the region's values in (lo, hi), in order (*n_valid of them), summed
 *      per 8192-value chunk in numpy's order with their logs beside
 *      (chunk_sums, 2 per chunk);
 *   3: out = amp - ALP (1 + log(amp / ALP)) where amp = flat * alp / avr >= ALP,
 *      else 0; count[1] = #(out[nL .. n - nL) > 1e-5); sorted = the window of
 *      out, ascending;
 *   4: out *= scale.
 * work: at least pycwb_g2g_work_bytes.  n < 2^31 - 1. */
typedef struct pycwb_g2g_desc {
  const double* flat;
  int64_t n;
  int64_t nL;
  int64_t nR;
  void* work;
  size_t work_bytes;
  uint64_t* count;
  double* sorted;
  double lo;
  double hi;
  double* chunk_sums;
  int64_t* n_valid;
  double alp;
  double avr;
  double ALP;
  double scale;
  double* out;
} pycwb_g2g_desc;

int pycwb_g2g_work_bytes(const pycwb_g2g_desc* desc, size_t* bytes);
int pycwb_g2g_phase(const pycwb_g2g_desc* desc, int phase, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PYCWB_AMD_H */
