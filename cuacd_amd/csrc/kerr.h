/* kerr.h - kernel error word: single-bit flags OR-ed into one int.
 *
 * Shared by the HIP kernels and the host C-ABI layer. Every kernel that can
 * fail ORs one of these bits into the caller's error word and the host turns
 * the word into text with kerr_decode(). The flag names follow the reference's
 * error vocabulary (cuda/error_codes.cuh:14-53) so error messages read the
 * same; the bit assignments are this library's own.
 */
#ifndef CUACD_KERR_H
#define CUACD_KERR_H

/* convex hull (wave quickhull, chunked divide-and-conquer) */
#define KERR_HULL_CAPACITY     (1 << 0)  /* hull larger than the workspace chosen for it   */
#define KERR_HULL_HEAP_OUTPUT  (1 << 1)  /* heap_alloc for the output hull mesh failed     */
#define KERR_HULL_TOPOLOGY     (1 << 2)  /* horizon / adjacency invariant broken            */
/* plane cut */
#define KERR_PC_SCRATCH_OOM    (1 << 3)
#define KERR_PC_POOL_OOM       (1 << 4)
#define KERR_PC_CAPACITY       (1 << 5)
/* Hausdorff */
#define KERR_HD_SCRATCH_OOM    (1 << 6)
#define KERR_HD_CAPACITY       (1 << 7)
/* connected components */
#define KERR_DC_OOM            (1 << 8)
/* k-DOP prefiltered hull */
#define KERR_KDOP_SCRATCH_OOM  (1 << 9)
/* lookahead search */
#define KERR_LA_OVERFLOW       (1 << 10)
#define KERR_LA_FRAME_OOM      (1 << 11)
#define KERR_LA_EVAL           (1 << 12)
/* mesh validation */
#define KERR_MESH_INVALID      (1 << 13)
/* device heap */
#define KERR_HEAP_OOM          (1 << 14)
#define KERR_HEAP_CORRUPT      (1 << 15)
/* host-injected fault (tests of error propagation) */
#define KERR_INJECTED          (1 << 16)
/* learned proposal */
#define KERR_NN_SHAPE          (1 << 17)

#define KERR_COUNT    18
#define KERR_ALL_BITS ((1 << KERR_COUNT) - 1)

/* host-only helper (no __device__ attribute), visible to both compile passes */
#include <stdio.h>
/* Decode an error word into "GPU error 0x.. [name, name]". Returns chars written. */
static inline int kerr_decode(int err, char* buf, int bufsize) {
    static const struct { int bit; const char* name; } table[] = {
        { KERR_HULL_CAPACITY,    "hull:capacity" },
        { KERR_HULL_HEAP_OUTPUT, "hull:heap_output_oom" },
        { KERR_HULL_TOPOLOGY,    "hull:topology" },
        { KERR_PC_SCRATCH_OOM,   "plane_cut:scratch_oom" },
        { KERR_PC_POOL_OOM,      "plane_cut:pool_oom" },
        { KERR_PC_CAPACITY,      "plane_cut:capacity" },
        { KERR_HD_SCRATCH_OOM,   "hausdorff:scratch_oom" },
        { KERR_HD_CAPACITY,      "hausdorff:capacity" },
        { KERR_DC_OOM,           "postprocess:oom" },
        { KERR_KDOP_SCRATCH_OOM, "kdop_hull:scratch_oom" },
        { KERR_LA_OVERFLOW,      "lookahead:overflow" },
        { KERR_LA_FRAME_OOM,     "lookahead:frame_oom" },
        { KERR_LA_EVAL,          "lookahead:eval" },
        { KERR_MESH_INVALID,     "mesh_invalid" },
        { KERR_HEAP_OOM,         "heap:oom" },
        { KERR_HEAP_CORRUPT,     "heap:corrupt" },
        { KERR_INJECTED,         "injected" },
        { KERR_NN_SHAPE,         "nn:shape" },
    };
    int off = 0, first = 1;
    if (!err || bufsize <= 0) { if (bufsize > 0) buf[0] = 0; return 0; }
    off += snprintf(buf + off, (size_t)(bufsize - off), "GPU error 0x%x [", err);
    for (size_t i = 0; i < sizeof(table) / sizeof(table[0]) && off < bufsize; i++) {
        if (err & table[i].bit) {
            off += snprintf(buf + off, (size_t)(bufsize - off), "%s%s", first ? "" : ", ", table[i].name);
            first = 0;
        }
    }
    if ((err & ~KERR_ALL_BITS) && off < bufsize)
        off += snprintf(buf + off, (size_t)(bufsize - off), "%sunknown:0x%x", first ? "" : ", ",
                        err & ~KERR_ALL_BITS);
    if (off < bufsize) off += snprintf(buf + off, (size_t)(bufsize - off), "]");
    return off;
}

#endif /* CUACD_KERR_H */
