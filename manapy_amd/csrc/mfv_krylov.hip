// libmfv batched sparse solves: (sigma_i I + tau_i A_i) x_i = b_i for many independent
// instances i stored as one block-diagonal ELL matrix in 64-row slices (one wavefront per
// slice). The solve is mixed precision: BiCGStab with Jacobi preconditioning in FP32 (FP64
// dot products) inside FP64 iterative refinement, r = b - op(x) in FP64 until
// ||r|| <= tol ||b||.
//
// Instances advance independently through a per-instance state machine kept in device memory
// (status + scalars); every kernel runs a fixed persistent grid over the live-slice or
// live-instance list and acts only on rows whose instance is in the state the kernel serves.
// A tick (mfv_krylov_tick) is therefore valid for any mix of instances at any stage and can be
// replayed from a captured graph while instances join and leave.
//
// Per-instance reductions are two-level and deterministic: each wave reduces its slice in a
// fixed shuffle pattern into partial[slice], then one block per instance adds the partials of
// its slices in a fixed order relative to its first slice. An instance's arithmetic therefore
// does not depend on where it sits in the batch or on what else is in it.
#include "mfv_common.hpp"

namespace mfv {

struct Ell {
    int S, W, I;
    const int32_t *__restrict__ slice_inst;
    const int32_t *__restrict__ slice_w;
    const int32_t *__restrict__ cols;
    const double *__restrict__ vals;
    const float *__restrict__ vals32;
    const double *__restrict__ diag;
    const int32_t *__restrict__ inst_slice0;
    const int32_t *__restrict__ inst_nslices;
    const int32_t *__restrict__ live_slices;
    const int32_t *__restrict__ n_live_slices;
    const int32_t *__restrict__ live_inst;
    const int32_t *__restrict__ n_live_inst;
    const int16_t *__restrict__ cols16;
    const uint8_t *__restrict__ narrow;
};

struct Work {
    double *__restrict__ x;
    double *__restrict__ b;
    float *__restrict__ d;
    float *__restrict__ r;
    float *__restrict__ rh;
    float *__restrict__ p;
    float *__restrict__ v;
    float *__restrict__ yz;
    float *__restrict__ t;
    float *__restrict__ dinv;
    double *__restrict__ partial;
    double *__restrict__ scal;
    int32_t *__restrict__ status;
};

static Ell to_kernel(const mfv_ell_batch *e)
{
    return Ell{e->nb_slices,    e->width,       e->nb_inst,       e->slice_inst, e->slice_w, e->cols,
               e->vals,         e->vals32,      e->diag,          e->inst_slice0, e->inst_nslices,
               e->live_slices,  e->n_live_slices, e->live_inst,   e->n_live_inst, e->cols16, e->slice_narrow};
}

static Work to_kernel(const mfv_krylov_work *w)
{
    return Work{w->x, w->b, w->d, w->r, w->rh, w->p, w->v, w->yz, w->t, w->dinv, w->partial, w->scal, w->status};
}

__device__ __forceinline__ double *S(const Work &w, int inst) { return w.scal + (size_t)inst * MFV_KS_NSCAL; }

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Iterates the live slices whose instance is in state `want`; body(slice, inst, row).
template <class Body>
__device__ __forceinline__ void for_slices(const Ell &e, const Work &w, int want0, int want1, Body body)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    const int n = *e.n_live_slices;
    for (int i = blockIdx.x * wpb + wave; i < n; i += gridDim.x * wpb) {
        const int s = e.live_slices[i];
        const int inst = e.slice_inst[s];
        if (inst < 0) continue;
        const int st = w.status[inst];
        if (st != want0 && st != want1) continue;
        body(s, inst, s * 64 + lane, lane);
    }
}

// sum_j A32[row][j] * x[col_j] in FP32 with explicit fused multiply-adds. A narrow slice reads
// its columns as 16-bit offsets from the row (the branch is uniform over the wave: one slice).
__device__ __forceinline__ bool slice_is_narrow(const Ell &e, int s) { return e.narrow && e.narrow[s]; }

// Sum over the row's slots j < wdt of val[j] * x[col[j]], accumulated in j order. The slots are
// taken in groups (8, then 4, then 1) whose column, value and x loads are all issued before the
// group's multiply-adds, so a slot no longer waits two memory latencies of its own; the
// accumulation order, hence the bits, are those of the slot-by-slot loop.
template <int G, class Val, class Col, class X, class Acc>
__device__ __forceinline__ int ell_group(int j, int wdt, size_t base, const Val *__restrict__ vals, Col col,
                                         const X *__restrict__ x, Acc &acc)
{
    for (; j + G <= wdt; j += G) {
        Val v[G];
        int c[G];
#pragma unroll
        for (int u = 0; u < G; u++) {
            const size_t q = base + (size_t)(j + u) * 64;
            v[u] = vals[q];
            c[u] = col(q);
        }
        X xv[G];
#pragma unroll
        for (int u = 0; u < G; u++) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < G; u++) acc = fma(v[u], xv[u], acc);
    }
    return j;
}

template <class Val, class X, class Acc>
__device__ __forceinline__ Acc ell_row(const Ell &e, int s, int lane, const Val *__restrict__ vals,
                                       const X *__restrict__ x)
{
    const int wdt = e.slice_w[s];
    const size_t base = (size_t)s * e.W * 64 + lane;
    Acc acc = 0;
    if (slice_is_narrow(e, s)) {
        const int row = s * 64 + lane;
        auto col = [&](size_t q) { return row + e.cols16[q]; };
        int j = ell_group<8>(0, wdt, base, vals, col, x, acc);
        j = ell_group<4>(j, wdt, base, vals, col, x, acc);
        ell_group<1>(j, wdt, base, vals, col, x, acc);
    } else {
        auto col = [&](size_t q) { return e.cols[q]; };
        int j = ell_group<8>(0, wdt, base, vals, col, x, acc);
        j = ell_group<4>(j, wdt, base, vals, col, x, acc);
        ell_group<1>(j, wdt, base, vals, col, x, acc);
    }
    return acc;
}

__device__ __forceinline__ float ell_row32(const Ell &e, int s, int lane, const float *__restrict__ x)
{
    return ell_row<float, float, float>(e, s, lane, e.vals32, x);
}

__device__ __forceinline__ double ell_row64(const Ell &e, int s, int lane, const double *__restrict__ x)
{
    return ell_row<double, double, double>(e, s, lane, e.vals, x);
}

// One wave per slice: the slice is narrow when every used slot's column is within +-32767 of its
// row; then its 16-bit offsets are written. Slots past the slice width are never read.
__global__ void __launch_bounds__(256) k_ell_narrow(Ell e, int slice0, int nslices, int16_t *__restrict__ cols16,
                                                     uint8_t *__restrict__ narrow)
{
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= nslices) return;
    const int s = slice0 + i;
    const int wdt = e.slice_w[s];
    const int row = s * 64 + lane;
    const size_t base = (size_t)s * e.W * 64 + lane;
    int wide = 0;
    for (int j = 0; j < wdt; j++) {
        const int d = e.cols[base + (size_t)j * 64] - row;
        wide |= (d < -32767 || d > 32767) ? 1 : 0;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) wide |= __shfl_xor(wide, off, 64);
    if (!wide)
        for (int j = 0; j < wdt; j++) {
            const size_t q = base + (size_t)j * 64;
            cols16[q] = (int16_t)(e.cols[q] - row);
        }
    if (lane == 0) narrow[s] = wide ? 0 : 1;
}

// ---------------------------------------------------------------------------------------------
// refinement: x += d (state CORRECT), then r = b - op(x) (states START and CORRECT)

__global__ void __launch_bounds__(256) k_ir_correct(Ell e, Work w)
{
    for_slices(e, w, MFV_KS_CORRECT, MFV_KS_CORRECT,
               [&](int, int, int row, int) { w.x[row] += (double)w.d[row]; });
}

__global__ void __launch_bounds__(256) k_ir_residual(Ell e, Work w)
{
    for_slices(e, w, MFV_KS_START, MFV_KS_CORRECT, [&](int s, int inst, int row, int lane) {
        const double *sc = S(w, inst);
        const double sigma = sc[MFV_KS_SIGMA], tau = sc[MFV_KS_TAU];
        if (w.status[inst] == MFV_KS_START) {
            const double dg = sigma + tau * e.diag[row];
            w.dinv[row] = (dg != 0.) ? (float)(1. / dg) : 0.f;
        }
        const double ax = ell_row64(e, s, lane, w.x);
        const double bb = w.b[row];
        const double res = bb - (sigma * w.x[row] + tau * ax);
        const float r32 = (float)res;
        w.r[row] = r32;
        w.rh[row] = r32;
        w.d[row] = 0.f;
        w.p[row] = 0.f;
        w.v[row] = 0.f;
        const double s0 = wave_sum(res * res);
        const double s1 = wave_sum(bb * bb);
        const double s2 = wave_sum((double)r32 * (double)r32);
        if (lane == 0) {
            w.partial[4 * s + 0] = s0;
            w.partial[4 * s + 1] = s1;
            w.partial[4 * s + 2] = s2;
        }
    });
}

// Threads of a per-instance block: an instance's reduction runs in one block, so a wide block
// shortens the serial part of the largest instance's sum (1M rows = 15 625 slice partials).
constexpr int kInstBlock = 1024;

// Sums partial[slice][0..n) over the instance's slices into out[0..n) (thread 0 holds them).
// Each thread keeps four running sums over its slices (i, i + B, i + 2B, i + 3B, then the next
// 4B), added in that fixed order, then a tree over the block: the order depends only on the
// instance's own slice count.
template <int NSUM>
__device__ __forceinline__ void inst_reduce(const Ell &e, const Work &w, int inst, double *out)
{
    __shared__ double red[NSUM][kInstBlock];
    const int s0 = e.inst_slice0[inst], ns = e.inst_nslices[inst];
    double acc[4][NSUM];
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
        for (int k = 0; k < NSUM; k++) acc[q][k] = 0.;
    for (int i = threadIdx.x; i < ns; i += 4 * kInstBlock) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int j = i + q * kInstBlock;
            if (j < ns) {
#pragma unroll
                for (int k = 0; k < NSUM; k++) acc[q][k] += w.partial[4 * (size_t)(s0 + j) + k];
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NSUM; k++) red[k][threadIdx.x] = (acc[0][k] + acc[1][k]) + (acc[2][k] + acc[3][k]);
    __syncthreads();
    for (int h = kInstBlock / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
#pragma unroll
            for (int k = 0; k < NSUM; k++) red[k][threadIdx.x] += red[k][threadIdx.x + h];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < NSUM; k++) out[k] = red[k][0];
}

// Per-instance kernels: one block per live instance (grid-stride over the live list).
template <class Body>
__device__ __forceinline__ void for_insts(const Ell &e, const Work &w, int want0, int want1, Body body)
{
    const int n = *e.n_live_inst;
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int inst = e.live_inst[i];
        const int st = w.status[inst];
        __syncthreads();
        if (st != want0 && st != want1) continue;
        body(inst);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kInstBlock) k_ir_check(Ell e, Work w)
{
    for_insts(e, w, MFV_KS_START, MFV_KS_CORRECT, [&](int inst) {
        double sum[3];
        inst_reduce<3>(e, w, inst, sum);
        if (threadIdx.x == 0) {
            double *sc = S(w, inst);
            const double resn2 = sum[0], bn2 = sum[1];
            sc[MFV_KS_RESN2] = resn2;
            sc[MFV_KS_BN2] = bn2;
            if (w.status[inst] == MFV_KS_START) sc[MFV_KS_OUTER] = 0.;
            sc[MFV_KS_OUTER] += 1.;
            const double tol = sc[MFV_KS_TOL_OUT];
            int st;
            if (!isfinite(resn2) || !isfinite(bn2)) st = MFV_KS_FAILED;  // NaN/inf data: refinement cannot help
            else if (resn2 <= tol * tol * bn2) st = MFV_KS_DONE;
            else if (sc[MFV_KS_OUTER] > sc[MFV_KS_MAXIT_OUT]) st = MFV_KS_FAILED;
            else {
                sc[MFV_KS_RHO] = sum[2];
                sc[MFV_KS_R0N2] = sum[2];
                sc[MFV_KS_RN2] = sum[2];
                sc[MFV_KS_ALPHA] = 1.;
                sc[MFV_KS_OMEGA] = 1.;
                sc[MFV_KS_BETA] = 0.;
                sc[MFV_KS_IT] = 0.;
                st = MFV_KS_INNER;
            }
            w.status[inst] = st;
        }
    });
}

// ---------------------------------------------------------------------------------------------
// BiCGStab iteration (right Jacobi preconditioning, y = M^-1 p, z = M^-1 s), FP32 vectors

__global__ void __launch_bounds__(256) k_bicg_p(Ell e, Work w)
{
    for_slices(e, w, MFV_KS_INNER, MFV_KS_INNER, [&](int, int inst, int row, int) {
        const double *sc = S(w, inst);
        float p;
        if (sc[MFV_KS_IT] == 0.) {
            p = w.r[row];
        } else {
            const float beta = (float)sc[MFV_KS_BETA], omega = (float)sc[MFV_KS_OMEGA];
            p = fmaf(beta, fmaf(-omega, w.v[row], w.p[row]), w.r[row]);
        }
        w.p[row] = p;
        w.yz[row] = w.dinv[row] * p;
    });
}

// v = op(y); partial = rh . v
__global__ void __launch_bounds__(256) k_bicg_spmv_v(Ell e, Work w)
{
    for_slices(e, w, MFV_KS_INNER, MFV_KS_INNER, [&](int s, int inst, int row, int lane) {
        const double *sc = S(w, inst);
        const float sigma = (float)sc[MFV_KS_SIGMA], tau = (float)sc[MFV_KS_TAU];
        const float v = fmaf(sigma, w.yz[row], tau * ell_row32(e, s, lane, w.yz));
        w.v[row] = v;
        const double s0 = wave_sum((double)w.rh[row] * (double)v);
        if (lane == 0) w.partial[4 * s] = s0;
    });
}

__global__ void __launch_bounds__(kInstBlock) k_bicg_alpha(Ell e, Work w)
{
    for_insts(e, w, MFV_KS_INNER, MFV_KS_INNER, [&](int inst) {
        double sum[1];
        inst_reduce<1>(e, w, inst, sum);
        if (threadIdx.x == 0) {
            double *sc = S(w, inst);
            const double rv = sum[0];
            const double a = sc[MFV_KS_RHO] / rv;
            if (rv == 0. || !isfinite(a)) {
                w.status[inst] = MFV_KS_CORRECT;  // breakdown: refine with what d holds
            } else {
                sc[MFV_KS_ALPHA] = a;
            }
        }
    });
}

// s = r - alpha v (in place of r); d += alpha y; z = M^-1 s (in place of y)
__global__ void __launch_bounds__(256) k_bicg_s(Ell e, Work w)
{
    for_slices(e, w, MFV_KS_INNER, MFV_KS_INNER, [&](int, int inst, int row, int) {
        const float alpha = (float)S(w, inst)[MFV_KS_ALPHA];
        const float s = fmaf(-alpha, w.v[row], w.r[row]);
        w.r[row] = s;
        w.d[row] = fmaf(alpha, w.yz[row], w.d[row]);
        w.yz[row] = w.dinv[row] * s;
    });
}

// t = op(z); partial = (t . s, t . t)
__global__ void __launch_bounds__(256) k_bicg_spmv_t(Ell e, Work w)
{
    for_slices(e, w, MFV_KS_INNER, MFV_KS_INNER, [&](int s, int inst, int row, int lane) {
        const double *sc = S(w, inst);
        const float sigma = (float)sc[MFV_KS_SIGMA], tau = (float)sc[MFV_KS_TAU];
        const float t = fmaf(sigma, w.yz[row], tau * ell_row32(e, s, lane, w.yz));
        w.t[row] = t;
        const double s0 = wave_sum((double)t * (double)w.r[row]);
        const double s1 = wave_sum((double)t * (double)t);
        if (lane == 0) {
            w.partial[4 * s + 0] = s0;
            w.partial[4 * s + 1] = s1;
        }
    });
}

__global__ void __launch_bounds__(kInstBlock) k_bicg_omega(Ell e, Work w)
{
    for_insts(e, w, MFV_KS_INNER, MFV_KS_INNER, [&](int inst) {
        double sum[2];
        inst_reduce<2>(e, w, inst, sum);
        if (threadIdx.x == 0) {
            double *sc = S(w, inst);
            const double om = (sum[1] != 0.) ? sum[0] / sum[1] : 0.;
            sc[MFV_KS_OMEGA] = isfinite(om) ? om : 0.;
        }
    });
}

// d += omega z; r = s - omega t; partial = (r . r, rh . r)
__global__ void __launch_bounds__(256) k_bicg_x(Ell e, Work w)
{
    for_slices(e, w, MFV_KS_INNER, MFV_KS_INNER, [&](int s, int inst, int row, int lane) {
        const float omega = (float)S(w, inst)[MFV_KS_OMEGA];
        w.d[row] = fmaf(omega, w.yz[row], w.d[row]);
        const float r = fmaf(-omega, w.t[row], w.r[row]);
        w.r[row] = r;
        const double s0 = wave_sum((double)r * (double)r);
        const double s1 = wave_sum((double)w.rh[row] * (double)r);
        if (lane == 0) {
            w.partial[4 * s + 0] = s0;
            w.partial[4 * s + 1] = s1;
        }
    });
}

__global__ void __launch_bounds__(kInstBlock) k_bicg_rho(Ell e, Work w)
{
    for_insts(e, w, MFV_KS_INNER, MFV_KS_INNER, [&](int inst) {
        double sum[2];
        inst_reduce<2>(e, w, inst, sum);
        if (threadIdx.x == 0) {
            double *sc = S(w, inst);
            const double rn2 = sum[0], rho_new = sum[1];
            const double omega = sc[MFV_KS_OMEGA];
            sc[MFV_KS_RN2] = rn2;
            sc[MFV_KS_IT] += 1.;
            sc[MFV_KS_ITTOT] += 1.;
            const double tol = sc[MFV_KS_TOL_IN];
            const bool done = rn2 <= tol * tol * sc[MFV_KS_R0N2] || sc[MFV_KS_IT] >= sc[MFV_KS_MAXIT_IN] ||
                              rho_new == 0. || omega == 0.;
            if (done) {
                w.status[inst] = MFV_KS_CORRECT;
            } else {
                sc[MFV_KS_BETA] = (rho_new / sc[MFV_KS_RHO]) * (sc[MFV_KS_ALPHA] / omega);
                sc[MFV_KS_RHO] = rho_new;
            }
        }
    });
}

// y = A x (FP64) on the live slices, whatever their state
__global__ void __launch_bounds__(256) k_ell_spmv(Ell e, const double *__restrict__ x, double *__restrict__ y)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int n = *e.n_live_slices;
    for (int i = blockIdx.x * 4 + wave; i < n; i += gridDim.x * 4) {
        const int s = e.live_slices[i];
        y[s * 64 + lane] = ell_row64(e, s, lane, x);
    }
}

static int validate_ell(const mfv_ell_batch *e)
{
    MFV_REQUIRE(e != nullptr, "ELL batch descriptor is NULL");
    MFV_REQUIRE(e->nb_slices >= 1 && e->width >= 1 && e->nb_inst >= 1, "ELL batch sizes must be positive");
    MFV_REQUIRE((e->cols16 == nullptr) == (e->slice_narrow == nullptr), "cols16 and slice_narrow go together");
    MFV_REQUIRE((long long)e->nb_slices * 64 < (1LL << 31), "ELL batch rows exceed int32");
    MFV_REQUIRE(e->slice_inst && e->slice_w && e->cols && e->vals && e->vals32 && e->diag && e->inst_slice0 &&
                    e->inst_nslices && e->live_slices && e->n_live_slices && e->live_inst && e->n_live_inst,
                "ELL batch descriptor has NULL pointers");
    return MFV_OK;
}

static int validate_work(const mfv_krylov_work *w)
{
    MFV_REQUIRE(w != nullptr, "Krylov work descriptor is NULL");
    MFV_REQUIRE(w->x && w->b && w->d && w->r && w->rh && w->p && w->v && w->yz && w->t && w->dinv && w->partial &&
                    w->scal && w->status,
                "Krylov work descriptor has NULL pointers");
    return MFV_OK;
}

}  // namespace mfv

using namespace mfv;

static int check_tick(const mfv_ell_batch *e, const mfv_krylov_work *w, int n_iter, int grid_slices, int grid_inst)
{
    if (int rc = validate_ell(e)) return rc;
    if (int rc = validate_work(w)) return rc;
    MFV_REQUIRE(n_iter >= 0 && n_iter <= 4096, "n_iter %d out of [0, 4096]", n_iter);
    MFV_REQUIRE(grid_slices >= 1 && grid_slices <= 65536, "grid_slices %d out of [1, 65536]", grid_slices);
    MFV_REQUIRE(grid_inst >= 1 && grid_inst <= 65536, "grid_inst %d out of [1, 65536]", grid_inst);
    return MFV_OK;
}

static int tick_launches(const mfv_ell_batch *e, const mfv_krylov_work *w, int n_iter, int grid_slices,
                         int grid_inst, hipStream_t st)
{
    const Ell ek = to_kernel(e);
    const Work wk = to_kernel(w);
    const dim3 gs(grid_slices), gi(grid_inst), b(256), bi(kInstBlock);
    hipLaunchKernelGGL(k_ir_correct, gs, b, 0, st, ek, wk);
    hipLaunchKernelGGL(k_ir_residual, gs, b, 0, st, ek, wk);
    hipLaunchKernelGGL(k_ir_check, gi, bi, 0, st, ek, wk);
    for (int it = 0; it < n_iter; it++) {
        hipLaunchKernelGGL(k_bicg_p, gs, b, 0, st, ek, wk);
        hipLaunchKernelGGL(k_bicg_spmv_v, gs, b, 0, st, ek, wk);
        hipLaunchKernelGGL(k_bicg_alpha, gi, bi, 0, st, ek, wk);
        hipLaunchKernelGGL(k_bicg_s, gs, b, 0, st, ek, wk);
        hipLaunchKernelGGL(k_bicg_spmv_t, gs, b, 0, st, ek, wk);
        hipLaunchKernelGGL(k_bicg_omega, gi, bi, 0, st, ek, wk);
        hipLaunchKernelGGL(k_bicg_x, gs, b, 0, st, ek, wk);
        hipLaunchKernelGGL(k_bicg_rho, gi, bi, 0, st, ek, wk);
    }
    return check_launch("krylov tick");
}

extern "C" int mfv_krylov_tick(const mfv_ell_batch *e, const mfv_krylov_work *w, int n_iter, int grid_slices,
                               int grid_inst, void *stream)
{
    if (int rc = check_tick(e, w, n_iter, grid_slices, grid_inst)) return rc;
    return tick_launches(e, w, n_iter, grid_slices, grid_inst, as_stream(stream));
}

extern "C" int mfv_krylov_graph(const mfv_ell_batch *e, const mfv_krylov_work *w, int n_iter, int grid_slices,
                                int grid_inst, void **graph_out)
{
    if (int rc = check_tick(e, w, n_iter, grid_slices, grid_inst)) return rc;
    MFV_REQUIRE(graph_out, "graph_out is NULL");
    *graph_out = nullptr;
    MfvGraph *g = new MfvGraph();
    int rc = graph_begin(g);
    if (rc == MFV_OK) rc = graph_finish(g, tick_launches(e, w, n_iter, grid_slices, grid_inst, g->stream));
    if (rc != MFV_OK) {
        graph_free(g);
        return rc;
    }
    *graph_out = g;
    return MFV_OK;
}

extern "C" int mfv_ell_narrow_columns(const mfv_ell_batch *e, int slice0, int nslices, int16_t *cols16,
                                      uint8_t *slice_narrow, void *stream)
{
    if (int rc = validate_ell(e)) return rc;
    MFV_REQUIRE(cols16 && slice_narrow, "NULL array");
    MFV_REQUIRE(slice0 >= 0 && nslices >= 0 && (long long)slice0 + nslices <= e->nb_slices,
                "slices [%d, %d) outside the batch's %d", slice0, slice0 + nslices, e->nb_slices);
    if (nslices == 0) return MFV_OK;
    hipLaunchKernelGGL(k_ell_narrow, dim3((nslices + 3) / 4), dim3(256), 0, as_stream(stream), to_kernel(e), slice0,
                       nslices, cols16, slice_narrow);
    return check_launch("k_ell_narrow");
}

extern "C" int mfv_ell_spmv(const mfv_ell_batch *e, const double *x, double *y, int grid_slices, void *stream)
{
    if (int rc = validate_ell(e)) return rc;
    MFV_REQUIRE(x && y, "NULL array");
    MFV_REQUIRE(grid_slices >= 1 && grid_slices <= 65536, "grid_slices %d out of [1, 65536]", grid_slices);
    hipLaunchKernelGGL(k_ell_spmv, dim3(grid_slices), dim3(256), 0, as_stream(stream), to_kernel(e), x, y);
    return check_launch("k_ell_spmv");
}
