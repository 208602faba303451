// Shared device helpers for libmfv (gfx950, wave64).
//
// Arithmetic in every kernel follows the reference expression order term by term
// and the library is compiled with -ffp-contract=off: with IEEE-correct division and
// square root, a cell-centric gather that adds the face contributions in the
// reference's face order reproduces the reference's face-ordered scatter bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/mfv.h"

namespace mfv {

void set_error(const char *fmt, ...);
int check_launch(const char *what);

constexpr int kBlock = 256;

inline unsigned grid_for(int n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

#define MFV_REQUIRE(cond, ...)            \
    do {                                  \
        if (!(cond)) {                    \
            ::mfv::set_error(__VA_ARGS__); \
            return MFV_EINVAL;            \
        }                                 \
    } while (0)

// Per-face-slot values of one cell (MF <= 4) as native vectors: element s with a
// lane-dependent s compiles to a select chain, where an array would be put in scratch memory.
typedef float slot_f __attribute__((ext_vector_type(4)));
typedef double slot_d __attribute__((ext_vector_type(4)));
typedef int slot_i __attribute__((ext_vector_type(4)));

// Python min(a, b) / max(a, b): the second argument wins only on a strict comparison.
__device__ __forceinline__ double pmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double pmax(double a, double b) { return (b > a) ? b : a; }

// dt reductions. Every dt-producing kernel reduces its wavefront in registers and lane 0
// stores the wave's minimum (non-negative values; NaN and negative candidates ignored) at
// part[global wave index] -- no atomics: thousands of same-address atomics queue at the L2
// for longer than the kernel's own work. dt_finish then folds the partials into the target
// (out[0] = min(out[0], min part)) in one block, deterministically.
// `gid` is the thread's global (logical) index, so the wave's slot is gid / 64.
__device__ __forceinline__ void wave_min_part(double *__restrict__ part, size_t gid, double v)
{
    double m = (v >= 0.0) ? v : __longlong_as_double(0x7ff0000000000000LL);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(m, off, 64);
        m = (o < m) ? o : m;
    }
    if ((threadIdx.x & 63) == 0) part[gid >> 6] = m;
}

// Maximum of v over the wavefront, as a wave-uniform value (for branches every lane takes the
// same way).
// The first KM stencil handles of one item, and its stencil and ghost counts, from a tile
// stencil table stored row-major with a row stride kp (a multiple of 8 uint16) whose last entry
// holds the counts (stencil | ghost << 8): whole 16-byte vectors, two loads for 16 handles and
// the counts instead of one load per handle and two for the counts. Handles at j >= K are 0.
template <int KM>
__device__ __forceinline__ void stencil_row(const uint16_t *__restrict__ lst, size_t row, int kp, int K,
                                            int (&h)[KM], int &nn, int &ng)
{
    static_assert(KM % 8 == 0, "whole groups of eight handles");
    const uint4 *r = reinterpret_cast<const uint4 *>(lst + row * (size_t)kp);
    const int gm = (kp - 1) >> 3;  // the group whose last entry is the counts
    unsigned cnt = 0u;
#pragma unroll
    for (int g = 0; g < KM / 8; g++) {
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (8 * g < K || g == gm) v = r[g];  // a row holds kp >= 8 (g + 1) entries in both cases
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int j = 8 * g + 2 * k;
            h[j] = (j < K) ? (int)(w[k] & 0xffffu) : 0;
            h[j + 1] = (j + 1 < K) ? (int)(w[k] >> 16) : 0;
        }
        if (g == gm) cnt = v.w >> 16;
    }
    if (gm >= KM / 8) cnt = lst[row * (size_t)kp + kp - 1];
    nn = (int)(cnt & 0xffu);
    ng = (int)(cnt >> 8);
}

__device__ __forceinline__ int wave_max_int(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return __builtin_amdgcn_readfirstlane(v);
}

// A captured graph (mfv_graph_launch / mfv_graph_destroy) with the capture stream and the dt
// partials buffer it owns: the library's shared scratch may be reallocated by later calls,
// which must not move memory a graph still refers to.
struct MfvGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t stream = nullptr;
    double *part = nullptr;
};
void graph_free(MfvGraph *g);
// Starts a capture on a new stream of g; returns MFV_OK or MFV_EHIP (g is then freed by the
// caller). graph_finish ends it and instantiates the graph.
int graph_begin(MfvGraph *g);
int graph_finish(MfvGraph *g, int rc);

// Cells at positions e and e + T of a tile's ext order (its nc cells from c0, then
// ext[e0 .. e0 + ne)), with both table loads issued before either is used. Past the end,
// the second is the first again (writing it twice is harmless).
__device__ __forceinline__ int2 tile_ext_pair(const int32_t *__restrict__ ext, int c0, int nc, int e0, int ne,
                                              int e, int T)
{
    const int eb = (e + T < nc + ne) ? e + T : e;
    int2 c = make_int2(c0 + e, c0 + eb);
    if (ne > 0) {
        const int xa = ext[e0 + min(max(e - nc, 0), ne - 1)];
        const int xb = ext[e0 + min(max(eb - nc, 0), ne - 1)];
        c.x = (e < nc) ? c.x : xa;
        c.y = (eb < nc) ? c.y : xb;
    }
    return c;
}

// XCD-aware block order. The dispatcher hands consecutive workgroups to the 8 XCDs in
// turn and each XCD has its own L2, so with blockIdx order a cell's mesh neighbours (close
// in cell number) are mostly processed, and cached, on other XCDs. The remap gives XCD x
// (the blocks b with b % 8 == x) one contiguous range of logical blocks instead.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block()
{
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int per = nb / kXcds, rem = nb % kXcds, x = b % kXcds;
    return x * per + (x < rem ? x : rem) + b / kXcds;
}

// Waves of a launch of grid_for(n, block) blocks.
inline size_t waves_for(int n, int block = kBlock) { return (size_t)grid_for(n, block) * (size_t)(block / 64); }

// Library-owned scratch for the per-wave partials, one buffer per (current device, stream),
// grown on demand; work on one stream runs in order, so reuse does not race. NULL if the
// allocation fails (error set).
double *dt_scratch(size_t nwaves, hipStream_t st);

// out[0] = min(out[0], part[0..n)) on the stream (one block).
int dt_finish(const double *part, size_t n, double *out, hipStream_t st);

// dt ring of the fused steps: slot k%3 is this step's dt, (k+1)%3 receives the next dt,
// (k+2)%3 is reset for the step after (nothing reads or writes it during step k); slot 3
// accumulates the simulated time (t += dt of each step).
struct DtRing {
    const double *cur;
    double *next;
    double *spare;
    double *time;
};
inline DtRing dt_ring(double *ring, int step)
{
    int k = ((step % 3) + 3) % 3;
    return DtRing{ring + k, ring + (k + 1) % 3, ring + (k + 2) % 3, ring + 3};
}
__device__ __forceinline__ void ring_advance(const DtRing &r)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        r.spare[0] = 1e6;
        r.time[0] += r.cur[0];
    }
}

// Mesh descriptor passed by value to kernels.
struct Mesh {
    int C, F, MF, K, G;
    const int32_t *__restrict__ face_cells;
    const int32_t *__restrict__ face_name;
    const double *__restrict__ face_normal;
    const double *__restrict__ face_mesure;
    const double *__restrict__ face_center;
    const double *__restrict__ face_ghost;
    const double *__restrict__ cell_center;
    const double *__restrict__ cell_volume;
    const int32_t *__restrict__ cell_faces;
    const int32_t *__restrict__ cell_nfaces;
    const uint8_t *__restrict__ order_split;
    const uint8_t *__restrict__ order_sorted;
    const int32_t *__restrict__ lsq_cells;
    const int32_t *__restrict__ lsq_ncells;
    const int32_t *__restrict__ lsq_ghosts;
    const int32_t *__restrict__ lsq_nghosts;
};

inline Mesh to_kernel(const mfv_mesh *m)
{
    Mesh k;
    k.C = m->nb_cells;
    k.F = m->nb_faces;
    k.MF = m->max_cell_faces;
    k.K = m->lsq_width;
    k.G = m->ghost_width;
    k.face_cells = m->face_cells;
    k.face_name = m->face_name;
    k.face_normal = m->face_normal;
    k.face_mesure = m->face_mesure;
    k.face_center = m->face_center;
    k.face_ghost = m->face_ghost;
    k.cell_center = m->cell_center;
    k.cell_volume = m->cell_volume;
    k.cell_faces = m->cell_faces;
    k.cell_nfaces = m->cell_nfaces;
    k.order_split = m->order_split;
    k.order_sorted = m->order_sorted;
    k.lsq_cells = m->lsq_cells;
    k.lsq_ncells = m->lsq_ncells;
    k.lsq_ghosts = m->lsq_ghosts;
    k.lsq_nghosts = m->lsq_nghosts;
    return k;
}

int validate_mesh(const mfv_mesh *m, int need_lsq, int need_order_split, int need_order_sorted);

// Checks the last launch and, if dt_host != NULL, copies dt_dev[0] to it (synchronising).
int finish_dt(double *dt_dev, double *dt_host, void *stream);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace mfv
