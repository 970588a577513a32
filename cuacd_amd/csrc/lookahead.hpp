// lookahead.hpp - task records shared by the lookahead kernels and their host.
//
// The lookahead search (reference csrc/lookahead.c:144-1124 and the la_*
// kernels under cuda/) is driven from the host over all meshes of a batch at
// once: every phase is one launch over a flat task table that mixes the
// meshes, so a launch holds thousands of workgroups even when each mesh has
// only a few parts. Meshes live in device-heap chunks ([verts | tris], freed
// with heap_free on the chunk base); tasks carry raw device pointers.
#pragma once
#include "ctx.hpp"
#include "hull.hpp"
#include "hausdorff.hpp"

namespace cuacd {

constexpr int LA_MAX_COMP = 256;  // components kept per part (more merge into the last)

// Hull of a point set: FP64 volume, optionally the hull mesh in a heap chunk.
// Sets above 256 points go through the k-DOP prefilter first (reference
// cuda/kdop_hull.cuh:49-398), which also prunes sets above 1024 points.
struct LaHullTask {
    const float* pts;    // xyz
    int n;
    int emit;            // 1: write the hull mesh into a heap-0 chunk (hv, ht); 2: the same
                         // when the heap has room (no error when it has not)
    double mesh_vol;     // for pruning
    double prune_gap;    // prune when max(rough_vol - mesh_vol, 0) > prune_gap; >= 1e29 disables
    QFrame frame;        // quantization frame of the full set (k-DOP path)
    int use_frame;
    int nf;              // filtered point count (k-DOP path)
    float* filt;         // filtered points (scratch-heap chunk, freed by the final hull)
    // results
    double vol;
    int pruned;
    int err;
    float* hv;           // hull mesh chunk: verts, then tris at ht
    int* ht;
    int hnv, hnt;
    int ncorner;         // corners of the canonical hull (emitted or not)
};

// Full plane cut of one part (level 0 of the search): both closed halves.
struct LaCutTask {
    const float* v;
    const int* t;
    int nv, nt;
    float plane[4];
    // results
    float* pv; int* pt; int pnv, pnt;
    float* nvp; int* ntp; int nnv, nnt;
    double pvol, nvol;
    float plo[3], phi[3], nlo[3], nhi[3];
    int status;          // 0 cut, 1 cap failed, 2 one side empty, <0 error
};

// Leaf scoring of one cut (last level): exact half-space volumes of both
// sides and the point sets their hulls are built from (scratch-heap chunks).
struct LaScoreTask {
    const float* v;
    const int* t;
    int nv, nt;
    float plane[4];
    // results
    double pvol, nvol;   // half-space volumes (divergence origin on the plane)
    int npt, nnt;        // triangles (whole or split pieces) on each side; 0 = empty side
    float* ppts; int np_pts;  // point sets of the two sides: one heap chunk,
    float* npts; int nn_pts;  // freed through ppts (npts points into it)
};

// Connected components of one part (union-find over triangle vertices).
struct LaCompOut {
    float* v; int* t; int nv, nt;
    double svol;
    float lo[3], hi[3];
};
struct LaCompTask {
    const float* v;
    const int* t;
    int nv, nt;
    // results
    // 0: one piece, part unchanged; else out[0..n_comp-1] hold the pieces
    // (a single one when inner shells were dropped from two or more)
    int n_comp;
    LaCompOut* out;      // LA_MAX_COMP slots
};

// Concave-edge detection of one part: records of shared edges whose dihedral
// angle exceeds the threshold (reference cuda/la_concave.cu:48-363).
struct LaEdgeRec {
    int lo, hi, f0, f1;
    float n0[3], n1[3], mid[3];
    int pad;
};
struct LaEdgeTask {
    const float* v;
    const int* t;
    int nv, nt;
    LaEdgeRec* out;      // capacity cap records
    int cap;
    int count;           // result (may exceed cap: then records were dropped)
};

// bbox of a mesh's vertices
struct LaBoxTask {
    const float* v;
    int nv;
    float lo[3], hi[3];
};

// Concatenation of two meshes (or two point sets when the triangle arrays are
// null): vertices of the first then the second; triangles of the first, then
// those of the second offset by n1, each dropped when all three of its
// vertices lie within 1e-3 of `plane` (has_plane) - the merge pass's
// shared-plane filter (reference cuda/postprocess_merge.cu:264-330).
struct LaGatherTask {
    const float* v1;
    int n1;
    const float* v2;
    int n2;
    const int* t1;
    int m1;
    const int* t2;
    int m2;
    float plane[4];
    int has_plane;
    float* dv;
    int* dt;
    int nt_out;  // result
};

// ---- launchers (la_kernels.hip); all stream-ordered on ctx->stream --------------
int la_launch_gather(gpu_ctx_t ctx, LaGatherTask* d_tasks, int n);
// Segmented device copy: segment i copies bytes[i] (a multiple of 4) from
// src[i] to dst[i]; one workgroup per segment.
struct LaSeg {
    const void* src;
    void* dst;
    unsigned long long bytes;
};
int la_launch_segcopy(gpu_ctx_t ctx, const LaSeg* d_segs, int n);
int la_launch_hull(gpu_ctx_t ctx, LaHullTask* d_tasks, const LaHullTask* h_tasks, int n);
// LDS-resident plane cuts (plane_cut_block_lds) for meshes of up to
// cut_lds_verts() vertices; pc_lds_launch sizes a launch whose largest mesh
// has max_nv vertices (tasks above *lds_nv take the heap-scratch path) and
// raises the kernel's dynamic-LDS limit to *lds_bytes.
int cut_lds_verts();
int pc_lds_launch(gpu_ctx_t ctx, const void* kernel, int max_nv, int* lds_nv, size_t* lds_bytes);
// max_nv: the most vertices of any task's mesh (sizes the LDS-resident cut)
int la_launch_cut(gpu_ctx_t ctx, LaCutTask* d_tasks, int n, int max_nv);
int la_launch_score(gpu_ctx_t ctx, LaScoreTask* d_tasks, int n, int max_nv);
int la_launch_components(gpu_ctx_t ctx, LaCompTask* d_tasks, int n);
int la_launch_edges(gpu_ctx_t ctx, LaEdgeTask* d_tasks, int n, float cos_pos, float cos_neg, int pos_mode,
                    int neg_mode);
int la_launch_bbox(gpu_ctx_t ctx, LaBoxTask* d_tasks, int n);
// heap_free every non-null pointer of the list (device array)
int la_launch_free(gpu_ctx_t ctx, void** d_ptrs, int n);

// Device-sized hull call: the hulls of tasks[0, n) (tasks with n <= 0 are
// skipped) with the k-DOP prefilter, the size classification and every tier's
// task list on the device - no host readback, so the launches can be
// captured in a HIP graph. `work` (la_hull_dev_work_bytes) holds the counters
// and lists, `arena` (la_hull_dev_arena_bytes) the retry lists and global
// workspaces. A set still above chunk_above points after the prefilter is not
// hulled: it sets work counter LA_HDEV_OVER, and the caller must take the
// host-driven path (la_launch_hull) instead.
constexpr int LA_HDEV_TIERS = 9;
constexpr int LA_HDEV_BIG = 0;                       // counters: [0] sets to prefilter,
constexpr int LA_HDEV_OVER = 1 + LA_HDEV_TIERS;      // [1, 10) per tier, [10] oversized sets,
constexpr int LA_HDEV_SUBN = LA_HDEV_OVER + 1;       // [11] chunk tasks used,
constexpr int LA_HDEV_FAIL = LA_HDEV_SUBN + 1;       // [12] sets left unhulled (see below)
constexpr int LA_HDEV_NCOUNT = 16;
// Oversized sets (above chunk_above points after the prefilter) are hulled in
// chunks of half that size first, as la_launch_hull does, when the caller gives
// a chunk table: the chunk hulls run as a nested device-sized call, then each
// set becomes the union of its chunks' corners. A set whose union is still
// too large (la_launch_hull would chunk it again) or whose chunks do not fit
// the table is left unhulled and counted in counter LA_HDEV_FAIL; the caller
// must then take the host-driven path.
struct LaHullDevChunks {
    LaHullTask* sub = nullptr;  // chunk task table
    int cap = 0;                // its capacity
    char* work = nullptr;       // la_hull_dev_work_bytes(cap)
    char* arena = nullptr;      // la_hull_dev_arena_bytes(ctx, cap)
    size_t arena_bytes = 0;
};
size_t la_hull_dev_work_bytes(int n);
size_t la_hull_dev_arena_bytes(gpu_ctx_t ctx, int n);
// prefilter = false skips the k-DOP pass (the chunk hulls of oversized sets,
// which la_launch_hull also hulls unfiltered, keeping their set's frame).
int la_launch_hull_dev(gpu_ctx_t ctx, LaHullTask* d_tasks, int n, char* work, char* arena, size_t arena_bytes,
                       const LaHullDevChunks* chunks = nullptr, bool prefilter = true);

}  // namespace cuacd

int hd_launch_tasks(gpu_ctx_t ctx, const cuacd::HdTask* d_tasks, int n, char* d_ws);
// the same with the task count on the device (at most cap tasks, *dcount present)
int hd_launch_tasks_dev(gpu_ctx_t ctx, const cuacd::HdTask* d_tasks, int cap, const int* dcount, int chunks,
                        char* d_ws);
