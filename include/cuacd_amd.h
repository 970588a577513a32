/* cuacd_amd.h - C ABI of the MI355X (gfx950) implementation of cuacd's GPU path.
 *
 * Plain C: no HIP, CUDA or PyTorch types cross this boundary. Every function
 * below is one the reference's CPython extension binds (csrc/module.c:532-562)
 * with the same name, argument meaning, ownership and error behaviour, unless
 * it is marked "no reference counterpart".
 *
 * Conventions (reference csrc/heap.h, csrc/test.h, csrc/lookahead.h):
 *   - host pointers in, caller-allocated host output buffers out;
 *   - return 0 on success, a HIP error code, a KERR_* bitmask, or -1;
 *     the text is available from gpu_last_error(ctx);
 *   - calls are synchronous unless named *_dev (device-pointer, stream-ordered
 *     variants for callers whose data already lives in HBM).
 */
#ifndef CUACD_AMD_H
#define CUACD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpu_ctx* gpu_ctx_t;

/* ---- context and device heap (replaces csrc/heap.h:24-64) ---------------- */

/* replaces gpu_set_fatbin (csrc/heap.h:30): kernels are linked into this
 * library, so the blob is accepted and ignored. */
void gpu_set_fatbin(const void* data, size_t len);

/* replaces gpu_init (csrc/heap.h:32, csrc/heap.c:59-220). device < 0 keeps the
 * current device; pool_bytes = 0 reserves 80% of free HBM capped at 64 GiB. */
int gpu_init(gpu_ctx_t* ctx, int device_ordinal, size_t pool_bytes);
/* replaces gpu_destroy (csrc/heap.h:33) */
void gpu_destroy(gpu_ctx_t ctx);
/* replaces gpu_last_error (csrc/heap.h:35); ctx may be NULL after a failed init */
const char* gpu_last_error(gpu_ctx_t ctx);
/* replaces gpu_heap_compact (csrc/heap.h:40): size-class heaps never fragment
 * across classes, so this only synchronises. */
int gpu_heap_compact(gpu_ctx_t ctx);
/* replaces gpu_pool_usage (csrc/heap.h:46): bytes bumped from the pool (monotonic) */
size_t gpu_pool_usage(gpu_ctx_t ctx);
/* replaces gpu_heap_stats (csrc/heap.h:51): heap_outstanding, heap_allocs,
 * heap_frees, scratch_outstanding, scratch_allocs, scratch_frees */
int gpu_heap_stats(gpu_ctx_t ctx, unsigned long long out[6]);

/* result of lookahead_decompose (csrc/heap.h:66-83) */
struct gpu_part_result {
    float* verts;
    int* tris;
    int nv;
    int nt;
    float* hull_verts;
    int* hull_tris;
    int hull_nv;
    int hull_nt;
    float mesh_vol;
    float hull_vol;
    float hausdorff;
};
struct gpu_result {
    struct gpu_part_result* parts;
    int nparts;
};
/* replaces gpu_result_free (csrc/heap.h:85) */
void gpu_result_free(struct gpu_result* result);

/* ---- batch kernels (replaces csrc/test.h:18-132) --------------------------- */

/* replaces gpu_test_warp_sort (csrc/test.h:23): sorts int4 (x,y,z,index)
 * records of each sub-array by (y, x, z, index), in place. */
int gpu_test_warp_sort(gpu_ctx_t ctx, int* points, int total_pts, const int* offsets, int n_arrays);

/* replaces gpu_hull_dandc (csrc/test.h:35-48): exact hull per cloud. */
int gpu_hull_dandc(gpu_ctx_t ctx, const float* pts, int total_pts, const int* offsets, int n_hulls,
                   int max_pts_per_hull, int max_hull_verts, int max_hull_tris, float* out_verts,
                   int* out_tris, int* out_nv, int* out_nt, int* out_errors);

/* no reference counterpart: gpu_hull_dandc plus the FP64 divergence-theorem
 * volume of every hull (the reference computes it on the host in float32,
 * cuacd/__init__.py:222-231). out_volumes may be NULL. */
int gpu_hull_dandc_vol(gpu_ctx_t ctx, const float* pts, int total_pts, const int* offsets, int n_hulls,
                       int max_pts_per_hull, int max_hull_verts, int max_hull_tris, float* out_verts,
                       int* out_tris, int* out_nv, int* out_nt, int* out_errors, double* out_volumes);

/* replaces gpu_test_mesh_volume (csrc/test.h:53-59) */
int gpu_test_mesh_volume(gpu_ctx_t ctx, const float* verts, int n_verts, const int* tris, int n_tris,
                         float* out_volume);

/* replaces gpu_batch_mesh_volume (csrc/test.h:70-79); vert_offsets may be NULL */
int gpu_batch_mesh_volume(gpu_ctx_t ctx, const float* verts, int total_verts, const int* tris,
                          int total_tris, const int* tri_offsets, const int* vert_offsets, int n_meshes,
                          float* out_volumes);

/* replaces gpu_kdop_hull (csrc/test.h:84-97) */
int gpu_kdop_hull(gpu_ctx_t ctx, const float* pts, int total_pts, const int* offsets, int n_hulls,
                  int max_hull_verts, int max_hull_tris, float* out_verts, int* out_tris, int* out_nv,
                  int* out_nt, float* out_volumes, int* out_errors);

/* replaces gpu_test_hausdorff (csrc/test.h:102-108): bidirectional sampled
 * Hausdorff distance between a hull and a mesh. */
int gpu_test_hausdorff(gpu_ctx_t ctx, const float* hull_verts, int hull_nv, const int* hull_tris,
                       int hull_nt, const float* mesh_verts, int mesh_nv, const int* mesh_tris,
                       int mesh_nt, float* out_hausdorff);

/* replaces gpu_test_plane_cut (csrc/test.h:118-127) */
int gpu_test_plane_cut(gpu_ctx_t ctx, const float* vertices, int n_verts, const int* triangles,
                       int n_tris, float pa, float pb, float pc_n, float pd, float* out_pos_verts,
                       int out_pos_verts_cap, int* out_pos_tris, int out_pos_tris_cap,
                       float* out_neg_verts, int out_neg_verts_cap, int* out_neg_tris,
                       int out_neg_tris_cap, int* out_n_pv, int* out_n_pt, int* out_n_nv,
                       int* out_n_nt);

/* replaces gpu_mesh_audit (csrc/test.h:132-136): MAV_* verdict bitmask */
int gpu_mesh_audit(gpu_ctx_t ctx, const float* verts, int n_verts, const int* tris, int n_tris,
                   unsigned int* out_flags);

/* replaces gpu_test_postprocess_dc (csrc/postprocess.h:20-27) */
int gpu_test_postprocess_dc(gpu_ctx_t ctx, const float* verts, int nv, const int* tris, int nt,
                            int max_components, int max_verts_per, int max_tris_per,
                            float* out_verts, int* out_tris, int* out_nv, int* out_nt,
                            int* out_n_components);

/* replaces lookahead_decompose (csrc/lookahead.h:28-43) */
int lookahead_decompose(gpu_ctx_t ctx, const float* verts, int nv, const int* tris, int nt,
                        const float* hull_verts, int hull_nv, const int* hull_tris, int hull_nt,
                        int max_iters, int width, int width2, float threshold, int depth,
                        int quick_depth, int max_n_cutting, int verbose, int debug,
                        int decompose_components, int decompose_components_per_iter,
                        int n_concave_edges, float concave_eps, float concave_threshold,
                        int concave_iters, int merge_hulls, struct gpu_result* out);

/* Batched lookahead_decompose (no reference counterpart): n_meshes packed
 * meshes (vert_off/tri_off prefix arrays of n_meshes + 1 entries) and their
 * convex hulls (hull_vert_off/hull_tri_off), decomposed independently with the
 * same parameters in one search whose launches span all meshes. out[m] gets
 * mesh m's parts (free each with gpu_result_free); out_iters[m] (optional) the
 * iterations it ran; out_stats (optional, 96 doubles): [0, 27) per phase -
 * components, part hulls, Hausdorff, concave edges, level-0 cuts, leaf
 * scoring, item hulls, merge, learned proposal - the GPU milliseconds,
 * launches and algorithmic bytes; [27, 35) host wall milliseconds by section
 * - set-up, per-iteration refresh and selection, concave planes, level 0,
 * deeper levels, evaluate and apply, post-loop passes and merge, result
 * readback; [35, 77) per hull kernel class - one-workgroup hulls of <= 64,
 * 128, 256, 512, 1024 points, multi-wave hulls of <= 2048, 4096, 8192, 16384
 * points, k-DOP prefilter, full-capacity retries of the 512-, 1024-, 2048-
 * and 4096-point first passes - the milliseconds between HIP events around
 * each launch, launches and algorithmic bytes (points read, volumes written);
 * [77, 80) learned proposal: level-0 candidates kept, candidates ranked,
 * fallbacks (gpu_lookahead_set_proposal); [80, 88) per host section, the GPU
 * milliseconds of the phases inside it; [88, 96) per host section, the host
 * waits on the context's streams. */
int gpu_lookahead_batch(gpu_ctx_t ctx, int n_meshes, const float* verts, const int* vert_off, const int* tris,
                        const int* tri_off, const float* hull_verts, const int* hull_vert_off,
                        const int* hull_tris, const int* hull_tri_off, int max_iters, int width, int width2,
                        float threshold, int depth, int quick_depth, int max_n_cutting,
                        int decompose_components, int decompose_components_per_iter, int n_concave_edges,
                        float concave_eps, float concave_threshold, int concave_iters, int merge_hulls,
                        struct gpu_result* out, int* out_iters, double* out_stats);

/* replaces gpu_preprocess (csrc/preprocess.h:18-24). The PaMO remesh is not
 * part of this library: always returns -1 with an explanatory error. */
int gpu_preprocess(gpu_ctx_t ctx, const float* verts, int nv, const int* tris, int nt, int resolution,
                   float** out_verts, int* out_nv, int** out_tris, int* out_nt);

/* ---- device-resident variants and plumbing (no reference counterpart) ------
 * For callers (benchmarks, pipelines) whose inputs already live in HBM. All
 * run on the context's stream and return without synchronising. */
int gpu_dev_alloc(gpu_ctx_t ctx, size_t bytes, void** out_ptr);
int gpu_dev_free(gpu_ctx_t ctx, void* ptr);
int gpu_memcpy_h2d(gpu_ctx_t ctx, void* dst, const void* src, size_t bytes);
int gpu_memcpy_d2h(gpu_ctx_t ctx, void* dst, const void* src, size_t bytes);
int gpu_memset_dev(gpu_ctx_t ctx, void* dst, int value, size_t bytes);
int gpu_sync(gpu_ctx_t ctx);
/* HIP events on the context's stream around a timed region */
int gpu_timer_start(gpu_ctx_t ctx);
int gpu_timer_stop(gpu_ctx_t ctx, float* elapsed_ms);
/* device name / CU count for reports */
int gpu_device_info(gpu_ctx_t ctx, char* name, int name_len, int* n_cu, size_t* hbm_bytes);

/* device-pointer form of gpu_hull_dandc_vol */
int gpu_hull_dandc_dev(gpu_ctx_t ctx, const float* d_pts, const int* d_offsets, int n_hulls,
                       int max_pts_per_hull, int max_hull_verts, int max_hull_tris, float* d_verts,
                       int* d_tris, int* d_nv, int* d_nt, int* d_errors, double* d_volumes);

/* gpu_hull_dandc_dev without any allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory - code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc. Same
 * arguments and the same outputs, bit for bit; clouds of at most 16384
 * points (max_pts_per_hull above that is rejected with an error). The
 * workspaces of clouds above 1024 points come from a fixed 64 MiB device
 * array per GPU (held until the process ends), in passes of as many clouds
 * as fit; it calls no hipMalloc, hipFree or allocating library routine.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (other streams, other contexts on the same device) race. */
int gpu_hull_dandc_dev_noalloc(gpu_ctx_t ctx, const float* d_pts, const int* d_offsets, int n_hulls,
                               int max_pts_per_hull, int max_hull_verts, int max_hull_tris, float* d_verts,
                               int* d_tris, int* d_nv, int* d_nt, int* d_errors, double* d_volumes);

/* ---- batched directed Hausdorff (no reference counterpart) -----------------
 * For pairs (A_i, B_i): out_ab[i] = max over samples p of A_i of the distance
 * from p to B_i's surface, out_ba[i] the other way round. mode 0 samples like
 * the reference's Hausdorff (max(int(2000*area), 1000) samples,
 * cuda/hausdorff.cuh:380-558); mode 1 draws n_samples stratified
 * area-uniform samples. Triangle indices are relative to each mesh's own
 * vertex range. Optional sample dumps (mode 1): n_pairs*n_samples*3 floats. */
int gpu_hausdorff_batch(gpu_ctx_t ctx, const float* a_verts, const int* a_vert_off, const int* a_tris,
                        const int* a_tri_off, const float* b_verts, const int* b_vert_off, const int* b_tris,
                        const int* b_tri_off, int n_pairs, int mode, int n_samples, float* out_ab, float* out_ba,
                        float* out_samples_a, float* out_samples_b);

/* Batched plane cuts (no reference counterpart; the lookahead search cuts
 * many (mesh, plane) pairs per step, cuda/lookahead.cu:270-420). Meshes are
 * packed (vert_off/tri_off prefix arrays, n_meshes + 1 entries); cut i cuts
 * mesh mesh_of[i] by plane planes[4i..4i+3] = (a, b, c, d). Per cut: FP64
 * volumes of both closed halves, sizes {pos nv, pos nt, neg nv, neg nt} and a
 * status (0 cut, 1 cap failed - mesh returned whole on the positive side,
 * 2 one side empty). Host arrays in and out. */
int gpu_plane_cut_batch(gpu_ctx_t ctx, const float* verts, const int* vert_off, const int* tris,
                        const int* tri_off, int n_meshes, const int* mesh_of, const float* planes, int n_cuts,
                        double* out_vol_pos, double* out_vol_neg, int* out_sizes, int* out_status);

/* Batched exact cut scoring (no reference entry point: the per-candidate work
 * of the lookahead's level 0, cuda/la_expand.cu:65-424, with the stop cost of
 * cuda/la_common.cuh:21-34). Cut i cuts mesh mesh_of[i] (packed as in
 * gpu_plane_cut_batch) by planes[4i..4i+3]; both closed halves get their
 * exact hulls; a half whose volume cost 0.3*rv is under tau also gets its
 * Hausdorff distance H to its hull (reference sampler, both directions,
 * cuda/la_refine.cu:19-45); its cost is max(0.3*rv, H). out_score[i] = the
 * larger cost of the two halves, 1e30 when the cut fails or leaves a side
 * empty. out_detail (optional, 10 doubles per cut): status, mesh volume
 * pos/neg (clamped to the hull volume), hull volume pos/neg, rv cost pos/neg,
 * H pos/neg (-1 when not measured), 1 if scored. out_stats (optional, 48
 * doubles): GPU ms of the cut, hull and Hausdorff stages, algorithmic bytes of
 * the cut and hull stages, hull tasks, then per hull kernel class (the 14
 * classes of gpu_lookahead_batch's statistics) ms, launches, bytes. */
int gpu_score_cuts(gpu_ctx_t ctx, const float* verts, const int* vert_off, const int* tris, const int* tri_off,
                   int n_meshes, const int* mesh_of, const float* planes, int n_cuts, float tau, float* out_score,
                   double* out_detail, double* out_stats);

/* Score plan (no reference counterpart; the scorer of gpu_score_cuts as one
 * device-sized level): the meshes and planes of a fixed batch are uploaded
 * once, and every stage is sized on the device - the cut, the hull tasks of
 * the halves, the k-DOP prefilter, each hull tier's task list and count, the
 * Hausdorff task table and workspace offsets, the frees - so one replay is a
 * captured HIP graph with no host synchronisation inside. create runs the
 * batch once eagerly (gpu_score_cuts) to size the Hausdorff workspace, once
 * through the device-sized launches (its scores must equal the eager ones),
 * then captures. run replays the graph, reads the per-half results back once
 * and writes the same out_score / out_detail as gpu_score_cuts. A batch with a
 * half above one workgroup's hull size after the prefilter is refused at
 * create (gpu_score_cuts handles it). */
typedef struct score_plan* score_plan_t;
int gpu_score_plan_create(gpu_ctx_t ctx, const float* verts, const int* vert_off, const int* tris,
                          const int* tri_off, int n_meshes, const int* mesh_of, const float* planes, int n_cuts,
                          float tau, score_plan_t* plan);
int gpu_score_plan_run(gpu_ctx_t ctx, score_plan_t plan, float* out_score, double* out_detail);
void gpu_score_plan_destroy(gpu_ctx_t ctx, score_plan_t plan);

/* Launch-shape knobs of the hull tiers and the proposal network's inference
 * (no reference counterpart; the autotuner sets them). They select kernels
 * and capacities only - every setting computes the same canonical hulls and
 * the same scores. Process-wide. Keys: 0 wide_min
 * (wave-tier sets above it run four waves per hull; 0 disables), 1 vcap8
 * (first-pass vertex capacity in eighths, 1..8), 2 half_cap (reduced-capacity
 * first passes, 0/1), 3 kdop_min (k-DOP prefilter above this many points,
 * >= 4; default 512), 4 chunk_above (prefiltered sets above it are hulled in chunks of
 * half that size first, 1024..16384), 5 multi (several apexes per quickhull
 * round in the eight-wave LDS tiers, 0/1), 6 mlp_fused (the proposal
 * network's inference as one fused point-layer kernel, 0/1; same scores),
 * 7 ldsv8_vcap (first-pass vertex capacity of the eight-wave LDS tier,
 * 64..1024 in steps of 32; hulls with more corners are redone in HBM),
 * 8 ldsb8_vcap (vertex capacity of an LDS first pass for sets of 2049-4096
 * points, 0 = none, else 64..1024; default 1024), 9 cut_lds_verts (plane
 * cuts of meshes with up to this many vertices keep their per-vertex state
 * in LDS and write no per-triangle scratch; 0 = none, else 256..110592;
 * default 98304; the same halves either way), 10 hull_lpt (each hull tier's
 * sets launched largest first, 0/1; default 0), 11 hull_b4_split (percent
 * of the 1025-2048-point sets, the smallest, hulled by the HBM-workspace
 * kernel beside the LDS tier, 0..90; default 0), 12 hull_order (the
 * four-wave and wave tiers of a hull call wait for the first passes of the
 * eight-wave LDS tiers, which need a whole CU each, 0/1), 13 hull_w8 (first
 * passes of the 513-1024-point (bit 0) and 257-512-point (bit 1) wave tiers
 * on eight waves with multi-apex rounds; bits 2 and 3: the same for the
 * full-capacity retries of these tiers; 0..15), 14 hull_multi_hbm
 * (multi-apex rounds in the HBM-workspace tiers, 0/1).
 * Values are clamped to their ranges. */
int gpu_set_tuning(gpu_ctx_t ctx, int key, int value);
int gpu_get_tuning(gpu_ctx_t ctx, int key, int* value);

/* Learned cut proposal (no reference counterpart): a point-set network that
 * predicts the exact cut score of a part's candidate planes. Input: 256
 * area-stratified surface samples per part in the part's box frame (center,
 * half the largest extent; gpu_mlp_sample_parts computes them), planes as
 * (unit normal, offset) in the same frame, part_of ascending. BF16 MFMA
 * layers with FP32 accumulation; FP32 master weights and Adam state on the
 * device. Parameters: gpu_mlp_n_params() floats (padded layout, see mlp.hip);
 * params = NULL draws a seeded He-uniform initialisation. */
typedef struct gpu_mlp* gpu_mlp_t;
int gpu_mlp_n_params(void);
int gpu_mlp_create(gpu_ctx_t ctx, const float* params, int seed, gpu_mlp_t* out);
void gpu_mlp_destroy(gpu_ctx_t ctx, gpu_mlp_t m);
int gpu_mlp_get_params(gpu_ctx_t ctx, gpu_mlp_t m, float* out);
int gpu_mlp_forward(gpu_ctx_t ctx, gpu_mlp_t m, const float* points, int n_parts, const float* planes,
                    const int* part_of, int n_cand, float* out_scores);
/* One training step on the mean squared error of the predicted scores against
 * targets: forward, backward (gradients into out_grads when given), Adam
 * update with learning rate lr when apply != 0. */
int gpu_mlp_train_step(gpu_ctx_t ctx, gpu_mlp_t m, const float* points, int n_parts, const float* planes,
                       const int* part_of, int n_cand, const float* targets, float lr, int apply, float* out_loss,
                       float* out_grads);
int gpu_mlp_sample_parts(gpu_ctx_t ctx, const float* verts, const int* vert_off, const int* tris, const int* tri_off,
                         int n_parts, float* out_points);
/* The network's BF16 MFMA GEMM on host operands (C = A . B, FP32 out; element
 * strides sam/sak, sbk/sbn; a_len/b_len elements) - checks of its tile layout. */
int gpu_mlp_gemm_test(gpu_ctx_t ctx, int M, int N, int K, const unsigned short* A, long long sam, long long sak,
                      long long a_len, const unsigned short* B, long long sbk, long long sbn, long long b_len,
                      float* C);
/* The lookahead's learned proposal: with m set, level 0 of each search
 * iteration evaluates exactly only the top_k candidates per cutting part by
 * predicted score (all of them when the kept ones yield no cut). m = NULL
 * restores the plain search. */
int gpu_lookahead_set_proposal(gpu_ctx_t ctx, gpu_mlp_t m, int top_k);
/* Training record of the search (no reference counterpart): while on, every
 * iteration appends, per cutting part, its 256 surface samples (box frame) and
 * per level-0 candidate the plane in that frame and the candidate's path cost
 * (what the search minimises; 1e30 when the cut gave no leaf). record(ctx, 0)
 * discards it; size/read copy it out (pts n_parts*256*3, planes n_cand*4). */
int gpu_lookahead_record(gpu_ctx_t ctx, int on);
int gpu_lookahead_record_size(gpu_ctx_t ctx, int* n_parts, int* n_cand);
int gpu_lookahead_record_read(gpu_ctx_t ctx, float* pts, float* planes, int* part_of, float* targets);
/* Level-0 candidate planes of whole meshes exactly as the search builds them
 * (width axis cuts inside the box margins, then the planes of n_concave
 * sampled concave edges; reference cuda/la_axis_cut.cuh:10-19,
 * cuda/la_concave.cu:48-363): mesh m's first out_count[m] (<= cap) planes at
 * out_planes[4 (m cap + j)], their search index k at out_k[m cap + j]. */
int gpu_lookahead_candidates(gpu_ctx_t ctx, int n_meshes, const float* verts, const int* vert_off, const int* tris,
                             const int* tri_off, int width, float threshold, int n_concave, float concave_eps,
                             float concave_threshold, int cap, float* out_planes, int* out_k, int* out_count);

/* Plans: the same computation on device-resident meshes, prepared once (task
 * table, workspaces, and the launches captured as a HIP graph) and replayed
 * many times (stream-ordered). Outputs are SQUARED distances. */
typedef struct hd_plan* hd_plan_t;
int gpu_hausdorff_plan_create(gpu_ctx_t ctx, const float* d_a_verts, const int* d_a_tris, const float* d_b_verts,
                              const int* d_b_tris, const int* h_a_vert_off, const int* h_a_tri_off,
                              const int* h_b_vert_off, const int* h_b_tri_off, int n_pairs, int mode,
                              int n_samples, float* d_out_ab2, float* d_out_ba2, float* d_samples_a,
                              float* d_samples_b, hd_plan_t* plan);
int gpu_hausdorff_plan_run(gpu_ctx_t ctx, hd_plan_t plan);
void gpu_hausdorff_plan_destroy(gpu_ctx_t ctx, hd_plan_t plan);

/* Known-size device copy (bytes, a multiple of 4; reps launches of the
 * library's segmented copy kernel in 1 MiB segments) for calibrating HBM
 * traffic counters; *ms = GPU ms per copy. No reference counterpart. */
int gpu_copy_probe(gpu_ctx_t ctx, size_t bytes, int reps, float* ms);

/* sets KERR_INJECTED in the next kernel error word read back by the library
 * (tests of error propagation; no GPU fault is involved) */
void gpu_debug_inject_error(gpu_ctx_t ctx, int kerr_bits);

#ifdef __cplusplus
}
#endif

#endif /* CUACD_AMD_H */
